// C ABI of the device layer (include/gauxc_amd/device.h): context, resident
// data, chunk planning, work lists and the stage pipeline of one EXC/VXC
// evaluation.
//
// HBM layout (all resident until replaced):
//   shells      DevShell[nshells] of every molecule, molecule after molecule
//   px/py/pz/w  SoA points and weights of all tasks, in upload order
//   shell list, local->molecule basis map, local shell -> first function
//   work        per chunk: [density items, 32-wide | 64-wide]
//                          [SYR2K tiles,   32-wide | 64-wide]
//   qoff        queue tables of the persistent launches
//   P, VXC      per molecule nbf x nbf, column-major, ld = nbf, concatenated
//   P (FP32)    FP32 copy of P (mixed precision), same offsets
//   red         (w eps rho, w rho) per point block, molecule-major
// Chunk scratch (reused by every chunk): per task [5][nbe][npts_pad]
// (B, dB/dx, dB/dy, dB/dz, Z).  Tasks are grouped greedily into chunks that
// fit the arena; each chunk is one launch per stage and tile width, so every
// launch carries thousands of workgroups (or one persistent workgroup per
// queue).
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <numeric>
#include <queue>
#include <string>
#include <vector>

namespace xcq {

static thread_local std::string g_last_error;
void set_error(const std::string& m) { g_last_error = m; }

}  // namespace xcq

using namespace xcq;

namespace {
constexpr int kStages = 6;

template <typename T>
int dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) return XCQ_SUCCESS;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc of ") + std::to_string(n * sizeof(T)) +
              " bytes failed: " + hipGetErrorString(e));
    (void)hipGetLastError();
    return XCQ_ERROR_OUT_OF_MEMORY;
  }
  return XCQ_SUCCESS;
}
template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// The command processor hands workgroup i of a launch to XCD i % 8.  Work
// lists are built task by task, so neighbouring items share a task's B/Z
// panels and P_sub gathers; this permutation gives each XCD runs of kXcdRun
// consecutive items (one L2 then sees all of a run's reuse) while keeping
// the list's large-first order at the granularity of 8 * kXcdRun items.
// The final partial group keeps its natural order.
constexpr size_t kXcds = 8, kXcdRun = 16;
void xcd_interleave(std::vector<WorkItem>& v, size_t b, size_t e) {
  constexpr size_t G = kXcds * kXcdRun;
  const size_t n = e - b, full = n / G * G;
  if (full == 0) return;
  std::vector<WorkItem> tmp(v.begin() + b, v.begin() + b + full);
  for (size_t i = 0; i < full; ++i) {
    const size_t g = i / G, o = i % G, x = o / kXcdRun, q = o % kXcdRun;
    v[b + g * G + q * kXcds + x] = tmp[i];
  }
}

// Dynamic-queue layout of items [b, e) of v (item costs c): nx sub-lists
// (one per XCD, so a task's point blocks or tiles share that XCD's L2),
// filled task by task, largest predicted task cost first, each task going
// to the sub-list with the least cost so far (longest-processing-time-first
// across XCDs); inside a sub-list tasks keep that decreasing-cost order and
// a task's items stay consecutive.  Rewrites v[b, e) sub-list after
// sub-list and appends the nx + 1 sub-list offsets (relative to b) to qoff.
void xcd_queue_layout(std::vector<WorkItem>& v, size_t b, size_t e, const std::vector<double>& c,
                      int nx, std::vector<int32_t>& qoff) {
  struct Group {
    size_t first, count;
    double cost;
  };
  std::vector<Group> groups;
  for (size_t i = b; i < e; ++i) {
    if (groups.empty() || v[i].task != v[groups.back().first].task)
      groups.push_back(Group{i, 0, 0.0});
    groups.back().count++;
    groups.back().cost += c[i - b];
  }
  std::stable_sort(groups.begin(), groups.end(),
                   [](const Group& x, const Group& y) { return x.cost > y.cost; });
  using QE = std::pair<double, int>;  // (load, sub-list)
  std::priority_queue<QE, std::vector<QE>, std::greater<QE>> heap;
  for (int q = 0; q < nx; ++q) heap.push({0.0, q});
  std::vector<std::vector<WorkItem>> qs(nx);
  for (const auto& g : groups) {
    QE top = heap.top();
    heap.pop();
    for (size_t k = 0; k < g.count; ++k) qs[top.second].push_back(v[g.first + k]);
    top.first += g.cost;
    heap.push(top);
  }
  size_t pos = b;
  qoff.push_back(0);
  for (int q = 0; q < nx; ++q) {
    for (const auto& w : qs[q]) v[pos++] = w;
    qoff.push_back((int32_t)(pos - b));
  }
}

struct MolInfo {
  int32_t sh_begin = 0, nshells = 0, nbf = 0;
  int64_t p_off = 0;
  bool have_density = false;
};
}  // namespace

struct xcq_context_s {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t arena_bytes = 0;  // user limit (0 = derive from free memory)
  double fill = 0.9;
  // Caller-owned device buffer used as the chunk scratch arena
  // (xcq_context_create_mem); never freed by the context.
  void* user_arena = nullptr;
  size_t user_arena_bytes = 0;
  int wall_khz = 100000;

  // molecules and basis
  DevShell* d_shells = nullptr;
  DevShellC* d_shc = nullptr;
  int32_t nshells = 0;
  std::vector<int32_t> sh_size, sh_first, sh_mol;  // sh_first: molecule-local
  std::vector<MolInfo> mols;
  int32_t max_nbf = 0;
  int32_t max_l = 0;  // highest shell angular momentum
  int64_t p_total = 0;
  int32_t* d_mol_n = nullptr;
  int64_t* d_mol_off = nullptr;
  int64_t* d_mol_blk = nullptr;  // nmol + 1 block offsets into red

  // tasks
  int32_t ntasks = 0;
  int64_t npts_total = 0, total_pblocks = 0;
  std::vector<DevTask> h_tasks;
  DevTask* d_tasks = nullptr;
  double *d_px = nullptr, *d_py = nullptr, *d_pz = nullptr, *d_w = nullptr;
  int32_t *d_shell_list = nullptr, *d_bfmap = nullptr, *d_lsh = nullptr;

  // work lists
  std::vector<Chunk> chunks;
  std::vector<WorkItem> h_work;
  std::vector<int32_t> h_qoff;
  WorkItem* d_work = nullptr;
  int32_t* d_qoff = nullptr;
  unsigned long long* d_counters = nullptr;  // ticket counters of the dynamic queues
  int64_t ncounters = 0;
  int32_t dens32_max_nbe = 48, syr2k32_max_nbe = 96;  // tile policy
  int32_t dens_hiocc_max_nbe = 0;  // 64-wide density tiles at three workgroups per CU
  int32_t dens128_min_nbe = 1 << 30;  // 128-wide (mixed) density tiles from this nbe on
  // Mixed density stage's basis evaluation (bases up to l = 3): 1 = its
  // fold and Z evaluate B / grad B, 3 = its fold does and a separate launch
  // forms Z, 0 = both read the planes the collocation stage stores.
  int32_t dens_in_stage = 3;
  double vxc_drop = kDefaultVxcDrop;  // see xcq_set_vxc_drop_tolerance
  int32_t nqueues = 0;                                 // 0: one workgroup per item
  std::vector<double> cost_d, cost_s;                  // per-task predicted costs
  bool timing = false;
  int64_t* d_tstamp = nullptr;

  // density / results
  double* d_P = nullptr;
  float* d_Pf32 = nullptr;  // FP32 copy of P (mixed precision)
  bool pf32_valid = false;
  double* d_vxc = nullptr;
  double* d_red = nullptr;
  // UKS and gradient per-point buffers, [5][pblocks][256]: density sums
  // (s, z), Z coefficients (s, z), w eps rho (gradient).
  double* d_uks = nullptr;
  double* d_Pz = nullptr;   // UKS: P_z = P_a - P_b (single molecule)
  double* d_vxcz = nullptr; // UKS: VXC_z
  bool have_pz = false;
  bool last_uks = false;    // the last evaluation was a UKS one
  double* d_tP = nullptr;   // FXC contraction: trial density (single molecule)
  bool have_tp = false;
  double* d_tPz = nullptr;  // UKS FXC contraction: trial tP_z
  bool have_tpz = false;
  double* d_fxcu = nullptr; // UKS FXC: [6][pblocks][256] density sums and Z coefficients
  bool last_fxc_uks = false;  // the last evaluation was a UKS FXC contraction
  bool last_fxc = false;    // the last evaluation was an FXC contraction (in d_vxc)
  std::vector<double> sh_origin;  // shell centres (3 per shell), host copy
  // EXC gradient: atoms, shell -> atom map, nearest-neighbour distances, result.
  int32_t grad_natoms = 0;
  double* d_gxyz = nullptr;
  double* d_dnear = nullptr;
  int32_t* d_gnbr_idx = nullptr;  // neighbour lists of the gradient's weight part
  double* d_gnbr_r = nullptr;
  bool ssf_lists = true;          // xcq_set_ssf_neighbour_lists
  bool grad_lists = false;        // the last gradient's weight part uses the lists
  double* d_grad = nullptr;
  int32_t* d_shell_atom = nullptr;
  bool last_grad = false;   // the last evaluation was an EXC gradient
  double* d_zc = nullptr;   // Z coefficients per point block (in-stage mode 3), lazily
  int64_t pblocks = 0;      // point blocks of the uploaded tasks
  double* d_out2 = nullptr;
  double* d_corr = nullptr;  // [nmol][kCorrBlocks] partials of the P32 correction
  double* d_rpart = nullptr;  // [nmol][kRedBlocks][2] partials of EXC/NEL
  double* d_tmax = nullptr;   // bound table: one {float, float} per function granule
  int64_t tmax_total = 0;
  int last_want_vxc = 0;     // the last evaluation returns VXC
  int last_vxc_eval = 0;     // the last evaluation computed VXC

  // chunk scratch
  void* d_scratch = nullptr;
  size_t scratch_bytes = 0;

  // graph of the whole evaluation
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr;
  int gkey = -1;

  // timing
  std::vector<hipEvent_t> ev;  // [chunk][stage][2]
  hipEvent_t ev_total[2] = {nullptr, nullptr};
  bool last_graph = false;
  int64_t counters[4] = {0, 0, 0, 0};

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
    gkey = -1;
  }
  void free_tasks() {
    dfree(d_tasks);
    dfree(d_px);
    dfree(d_py);
    dfree(d_pz);
    dfree(d_w);
    dfree(d_shell_list);
    dfree(d_bfmap);
    dfree(d_lsh);
    dfree(d_work);
    dfree(d_qoff);
    dfree(d_counters);
    dfree(d_red);
    dfree(d_zc);
    dfree(d_uks);
    dfree(d_fxcu);
    pblocks = 0;
    dfree(d_tmax);
    tmax_total = 0;
    dfree(d_tstamp);
    if (d_scratch && d_scratch != user_arena) (void)hipFree(d_scratch);
    d_scratch = nullptr;
    scratch_bytes = 0;
    for (auto e : ev) (void)hipEventDestroy(e);
    ev.clear();
    chunks.clear();
    h_work.clear();
    h_qoff.clear();
    ntasks = 0;
  }
  ~xcq_context_s() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graph();
    free_tasks();
    for (auto e : ev_total)
      if (e) (void)hipEventDestroy(e);
    dfree(d_shells);
    dfree(d_shc);
    dfree(d_mol_n);
    dfree(d_mol_off);
    dfree(d_mol_blk);
    dfree(d_P);
    dfree(d_Pf32);
    dfree(d_vxc);
    dfree(d_Pz);
    dfree(d_vxcz);
    dfree(d_tP);
    dfree(d_tPz);
    dfree(d_gxyz);
    dfree(d_dnear);
    dfree(d_gnbr_idx);
    dfree(d_gnbr_r);
    dfree(d_grad);
    dfree(d_shell_atom);
    dfree(d_out2);
    dfree(d_corr);
    dfree(d_rpart);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

CtxView view_of(xcq_context_s* c) {
  return CtxView{c->stream,  c->d_tasks, c->d_shells, c->d_shell_list, c->d_bfmap,
                 c->d_lsh,   c->d_px,    c->d_py,     c->d_pz,         c->d_w,
                 c->d_work,  c->d_qoff,  c->d_counters, c->timing ? c->d_tstamp : nullptr,
                 c->d_P,     c->d_Pf32,  c->d_vxc,   c->d_red,        c->d_scratch,
                 c->vxc_drop, reinterpret_cast<unsigned*>(c->d_tmax),
                 c->max_l <= kInStageMaxL ? c->dens_in_stage : 0, c->d_shc, c->d_zc};
}

// Heuristic per-item costs (used when no predicted task costs are set):
// matrix-core work of the item, npts * rows * cols in the padded tile.
double heur_dens_item(const DevTask& d) {
  const double rows = std::ceil(d.nbe / double(d.nt_d)) * d.nt_d;
  return (double)kPtBlock * rows * d.nbe;
}
double heur_syr2k_item(const DevTask& d) { return (double)d.npts_pad * d.nt_s * d.nt_s; }

// Rebuild the work lists and queue tables of every chunk from the current
// tile policy and schedule, and upload them.
int build_work_lists(xcq_context_s* ctx) {
  auto& W = ctx->h_work;
  auto& Q = ctx->h_qoff;
  W.clear();
  Q.clear();
  int64_t ncounters = 0;
  const bool have_cost = ctx->cost_d.size() == (size_t)ctx->ntasks &&
                         ctx->cost_s.size() == (size_t)ctx->ntasks;
  for (auto& d : ctx->h_tasks) {
    d.nt_d = d.nbe <= ctx->dens32_max_nbe ? 32 : (d.nbe >= ctx->dens128_min_nbe ? 128 : 64);
    d.dens_cls = d.nt_d == 32    ? 0
                 : d.nt_d == 128 ? 3
                                 : (d.nbe <= ctx->dens_hiocc_max_nbe ? 1 : 2);
    d.nt_s = d.nbe <= ctx->syr2k32_max_nbe ? 32 : 64;
  }
  auto finish_range = [&](WorkRange& r, const std::vector<double>& cost) {
    r.n = (int64_t)W.size() - r.off;
    r.nq = 0;
    r.qoff_off = (int64_t)Q.size();
    if (r.n == 0) return;
    if (ctx->nqueues > 0) {
      r.nq = (int32_t)std::min<int64_t>(ctx->nqueues, r.n);
      // Workgroup b runs on XCD b % 8 (round-robin dispatch): eight
      // sub-lists when every XCD gets a workgroup, else one shared list.
      r.nx = r.nq >= (int32_t)kXcds ? (int32_t)kXcds : 1;
      r.cnt_off = ncounters;
      ncounters += r.nx;
      xcd_queue_layout(W, (size_t)r.off, W.size(), cost, r.nx, Q);
    } else {
      xcd_interleave(W, (size_t)r.off, W.size());
    }
  };
  for (auto& c : ctx->chunks) {
    for (int v = 0; v < 4; ++v) {  // density classes (DevTask::dens_cls)
      c.dens[v].off = (int64_t)W.size();
      std::vector<double> cost;
      for (int32_t i = c.task_begin; i < c.task_end; ++i) {
        const DevTask& d = ctx->h_tasks[i];
        if (d.dens_cls != v) continue;
        const int nmb = d.npts_pad / kPtBlock;
        const double ci = have_cost ? ctx->cost_d[i] / nmb : heur_dens_item(d);
        for (int pb = 0; pb < nmb; ++pb) {
          W.push_back(WorkItem{i, pb, 0, 0});
          cost.push_back(ci);
        }
      }
      finish_range(c.dens[v], cost);
    }
    c.coll.off = c.dens[0].off;
    c.coll.n = c.dens[0].n + c.dens[1].n + c.dens[2].n + c.dens[3].n;
    c.coll.nq = 0;
    for (int v = 0; v < 2; ++v) {  // SYR2K tiles
      const int nt = v == 0 ? 32 : 64;
      c.syr2k[v].off = (int64_t)W.size();
      std::vector<double> cost;
      for (int32_t i = c.task_begin; i < c.task_end; ++i) {
        const DevTask& d = ctx->h_tasks[i];
        if (d.nt_s != nt) continue;
        const int nt_tiles = (d.nbe + nt - 1) / nt;
        const int npairs = nt_tiles * (nt_tiles + 1) / 2;
        const double ci = have_cost ? ctx->cost_s[i] / npairs : heur_syr2k_item(d);
        for (int it = 0; it < nt_tiles; ++it)
          for (int jt = 0; jt <= it; ++jt) {
            W.push_back(WorkItem{i, it, jt, 0});
            cost.push_back(ci);
          }
      }
      finish_range(c.syr2k[v], cost);
    }
  }
  dfree(ctx->d_work);
  dfree(ctx->d_qoff);
  dfree(ctx->d_tstamp);
  dfree(ctx->d_counters);
  int rc;
  if ((rc = dalloc(&ctx->d_work, W.size()))) return rc;
  if ((rc = dalloc(&ctx->d_qoff, Q.size()))) return rc;
  ctx->ncounters = ncounters;
  if (ncounters && (rc = dalloc(&ctx->d_counters, (size_t)ncounters))) return rc;
  if (ctx->timing && (rc = dalloc(&ctx->d_tstamp, 2 * W.size()))) return rc;
  if (!W.empty())
    XCQ_HIP(hipMemcpy(ctx->d_work, W.data(), W.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
  if (!Q.empty())
    XCQ_HIP(hipMemcpy(ctx->d_qoff, Q.data(), Q.size() * 4, hipMemcpyHostToDevice));
  if (ctx->d_tstamp)
    XCQ_HIP(hipMemset(ctx->d_tstamp, 0, 2 * W.size() * sizeof(int64_t)));
  ctx->drop_graph();
  return XCQ_SUCCESS;
}

// Enqueue one evaluation on the context's stream.  Events are recorded only
// when `events` is set (not under graph capture).
// The basis evaluation mode a precision runs: the FP64 kernels have the
// stored planes (0) and the in-stage fold with a Z launch (3).
int effective_mode(const CtxView& v, int precision) {
  if (precision == XCQ_PREC_FP64) return v.in_stage != 0 ? 3 : 0;
  return v.in_stage;
}

int enqueue_pipeline(xcq_context_s* ctx, int func, int precision, int want_vxc, bool events) {
  const int gga = (func == XCQ_FUNC_PBE || func == XCQ_FUNC_PBE0) ? 1 : 0;
  const CtxView v = view_of(ctx);
  const int mode = effective_mode(v, precision);
  hipStream_t s = ctx->stream;
  const double fac = func == 0 ? 1.0 : 2.0;
  int rc;
  if (want_vxc && (rc = launch_fill(s, ctx->d_vxc, ctx->p_total))) return rc;
  if (want_vxc && (rc = launch_fill(s, ctx->d_tmax, ctx->tmax_total))) return rc;
  // Ticket counters of the dynamic queues start at zero (a fill kernel, not
  // a memset node, so graph replays reset them too; zero has the same bits
  // as 0.0).
  if (ctx->ncounters &&
      (rc = launch_fill(s, reinterpret_cast<double*>(ctx->d_counters), ctx->ncounters)))
    return rc;
  int64_t launches = want_vxc ? 1 : 0, items = 0;
  auto rec = [&](hipEvent_t e) { return events ? hipEventRecord(e, s) : hipSuccess; };
  for (size_t ci = 0; ci < ctx->chunks.size(); ++ci) {
    const Chunk& c = ctx->chunks[ci];
    hipEvent_t* ev = ctx->ev.data() + ci * kStages * 2;
    XCQ_HIP(rec(ev[0]));
    const WorkItem* cw = v.work + c.coll.off;
    if (precision == XCQ_PREC_FP64)
      rc = launch_collocation_fp64(s, cw, c.coll.n, v.tasks, v.shells, v.shell_list, v.lsh, v.px,
                                   v.py, v.pz, static_cast<double*>(v.scratch),
                                   gga && mode == 0, ctx->max_l);
    else
      rc = launch_collocation_fp32(s, cw, c.coll.n, v.tasks, v.shells, v.shell_list, v.lsh, v.px,
                                   v.py, v.pz, static_cast<float*>(v.scratch),
                                   gga && mode == 0, ctx->max_l);
    if (rc) return rc;
    XCQ_HIP(rec(ev[1]));
    // X, rho/grad rho, functional, EXC/NEL and Z in one launch per tile
    // width (stage 1); stage 2 (functional) stays an empty interval so the
    // stage table keeps its shape, and so does stage 3 (Z) except in the
    // mixed path's in-stage mode 3, where it times the Z launch.
    XCQ_HIP(rec(ev[2]));
    for (int k = 0; k < 4; ++k)
      if ((rc = launch_density(v, c.dens[k], k == 0 ? 32 : (k == 3 ? 128 : 64), k == 1,
                               precision, func, gga, fac, want_vxc)))
        return rc;
    XCQ_HIP(rec(ev[3]));
    XCQ_HIP(rec(ev[4]));
    XCQ_HIP(rec(ev[5]));
    launches += 5;
    items += c.coll.n;
    if (want_vxc) {
      XCQ_HIP(rec(ev[6]));
      // In-stage mode 3: Z (stage 3) from the coefficients the density
      // stage left.
      if (mode == 3) {
        if ((rc = launch_z_in_stage(s, cw, c.coll.n, v.tasks, v.shells, v.shell_list, v.lsh,
                                    v.px, v.py, v.pz, v.zc, v.scratch, precision,
                                    v.vxc_drop > 0.0 ? v.tmax : nullptr, gga, ctx->max_l)))
          return rc;
        ++launches;
      }
      XCQ_HIP(rec(ev[7]));
      XCQ_HIP(rec(ev[8]));
      for (int k = 0; k < 2; ++k)
        if ((rc = launch_syr2k(v, c.syr2k[k], k == 0 ? 32 : 64, precision))) return rc;
      XCQ_HIP(rec(ev[9]));
      launches += 2;
      items += c.syr2k[0].n + c.syr2k[1].n;
    }
  }
  const int32_t nmol = (int32_t)ctx->mols.size();
  // Mixed precision: EXC gets the first-order correction for the FP32
  // rounding of P, fac * sum (P - P32) . VXC (run_pipeline always asks for
  // VXC in that precision).
  const bool corr = precision == XCQ_PREC_MIXED && want_vxc && func != 0;
  if (want_vxc) {
    XCQ_HIP(rec(ctx->ev[10]));
    if ((rc = launch_symmetrize(s, nmol, ctx->max_nbf, ctx->d_mol_n, ctx->d_mol_off,
                                ctx->d_vxc)))
      return rc;
    if (corr && (rc = launch_p32_correction(s, nmol, ctx->d_mol_n, ctx->d_mol_off, ctx->d_P,
                                            ctx->d_Pf32, ctx->d_vxc, ctx->d_corr)))
      return rc;
    XCQ_HIP(rec(ctx->ev[11]));
    launches += corr ? 2 : 1;
  }
  if ((rc = launch_reduce_sums(s, nmol, ctx->d_mol_blk, ctx->d_red, ctx->d_rpart,
                               corr ? ctx->d_corr : nullptr, fac, ctx->d_out2)))
    return rc;
  ctx->counters[0] = launches + 1;
  ctx->counters[1] = items;
  return XCQ_SUCCESS;
}

int run_pipeline(xcq_context_s* ctx, int func, int precision, int want_vxc) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(ctx->ntasks > 0, XCQ_ERROR_STATE, "no tasks uploaded");
  for (const auto& m : ctx->mols)
    XCQ_REQUIRE(m.have_density, XCQ_ERROR_STATE, "no density uploaded");
  XCQ_REQUIRE(func == 0 || func == XCQ_FUNC_SVWN5 || func == XCQ_FUNC_PBE || func == XCQ_FUNC_PBE0,
              XCQ_ERROR_INVALID_ARGUMENT, "unknown functional");
  XCQ_REQUIRE(precision == XCQ_PREC_FP64 || precision == XCQ_PREC_MIXED,
              XCQ_ERROR_INVALID_ARGUMENT, "unknown precision");
  XCQ_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  if (precision == XCQ_PREC_MIXED && !ctx->pf32_valid) {
    if (!ctx->d_Pf32) {
      int rc = dalloc(&ctx->d_Pf32, (size_t)ctx->p_total);
      if (rc) return rc;
      ctx->drop_graph();
    }
    int rc = launch_P_to_f32(s, ctx->p_total, ctx->d_P, ctx->d_Pf32);
    if (rc) return rc;
    ctx->pf32_valid = true;
  }
  ctx->last_want_vxc = want_vxc;
  ctx->last_uks = false;
  ctx->last_fxc = false;
  ctx->last_fxc_uks = false;
  ctx->last_grad = false;
  // The mixed path's EXC correction needs VXC, so EXC-only requests in that
  // precision evaluate it too (and do not return it).
  if (precision == XCQ_PREC_MIXED && func != 0) want_vxc = 1;
  if (want_vxc && effective_mode(view_of(ctx), precision) == 3 && !ctx->d_zc) {
    int rc = dalloc(&ctx->d_zc, (size_t)256 * std::max<int64_t>(ctx->pblocks, 1));
    if (rc) return rc;
    ctx->drop_graph();
  }
  ctx->last_vxc_eval = want_vxc;
  if (!ctx->use_graph) {
    ctx->last_graph = false;
    return enqueue_pipeline(ctx, func, precision, want_vxc, true);
  }
  // Graph mode: capture on first use of this (functional, precision,
  // want_vxc) and replay afterwards.
  const int key = (func * 4 + precision) * 2 + want_vxc;
  if (!ctx->gexec || ctx->gkey != key) {
    ctx->drop_graph();
    XCQ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_pipeline(ctx, func, precision, want_vxc, false);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    XCQ_HIP(e);
    e = hipGraphInstantiate(&ctx->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    XCQ_HIP(e);
    ctx->gkey = key;
  }
  XCQ_HIP(hipEventRecord(ctx->ev_total[0], s));
  XCQ_HIP(hipGraphLaunch(ctx->gexec, s));
  XCQ_HIP(hipEventRecord(ctx->ev_total[1], s));
  ctx->last_graph = true;
  return XCQ_SUCCESS;
}

}  // namespace

extern "C" {

const char* xcq_last_error(void) { return g_last_error.c_str(); }

int xcq_device_count(int* count) {
  XCQ_REQUIRE(count, XCQ_ERROR_INVALID_ARGUMENT, "null count");
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    (void)hipGetLastError();
    set_error(std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
    return XCQ_ERROR_NO_DEVICE;
  }
  return XCQ_SUCCESS;
}

int xcq_context_create(int device, size_t arena_bytes, double fill_fraction, xcq_context* out) {
  XCQ_REQUIRE(out, XCQ_ERROR_INVALID_ARGUMENT, "null context pointer");
  *out = nullptr;
  int n = 0;
  if (xcq_device_count(&n) != XCQ_SUCCESS || n == 0) {
    set_error("no HIP device available");
    return XCQ_ERROR_NO_DEVICE;
  }
  XCQ_REQUIRE(device >= 0 && device < n, XCQ_ERROR_INVALID_ARGUMENT, "device out of range");
  XCQ_REQUIRE(fill_fraction > 0. && fill_fraction <= 1., XCQ_ERROR_INVALID_ARGUMENT,
              "fill_fraction must be in (0, 1]");
  XCQ_HIP(hipSetDevice(device));
  auto ctx = std::make_unique<xcq_context_s>();
  ctx->device = device;
  ctx->arena_bytes = arena_bytes;
  ctx->fill = fill_fraction;
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess &&
      khz > 0)
    ctx->wall_khz = khz;
  else
    (void)hipGetLastError();
  XCQ_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  XCQ_HIP(hipEventCreate(&ctx->ev_total[0]));
  XCQ_HIP(hipEventCreate(&ctx->ev_total[1]));
  int rc = upload_sph_table(ctx->stream);
  if (rc) return rc;
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  *out = ctx.release();
  return XCQ_SUCCESS;
}

int xcq_context_create_mem(int device, void* mem, size_t mem_bytes, xcq_context* out) {
  XCQ_REQUIRE(out, XCQ_ERROR_INVALID_ARGUMENT, "null context pointer");
  *out = nullptr;
  XCQ_REQUIRE(mem || mem_bytes == 0, XCQ_ERROR_INVALID_ARGUMENT,
              "null device buffer with a non-zero size");
  XCQ_REQUIRE(((uintptr_t)mem & 255) == 0, XCQ_ERROR_INVALID_ARGUMENT,
              "the device buffer must be 256-byte aligned");
  xcq_context ctx = nullptr;
  int rc = xcq_context_create(device, mem_bytes ? mem_bytes : 1, 0.9, &ctx);
  if (rc) return rc;
  ctx->user_arena = mem;
  ctx->user_arena_bytes = mem_bytes;
  *out = ctx;
  return XCQ_SUCCESS;
}

int xcq_context_destroy(xcq_context ctx) {
  delete ctx;
  return XCQ_SUCCESS;
}

int xcq_synchronize(xcq_context ctx) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_HIP(hipSetDevice(ctx->device));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  return XCQ_SUCCESS;
}

int xcq_upload_molecules(xcq_context ctx, int32_t nmol, const int32_t* mol_nshells,
                         const xcq_shell* shells) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(nmol > 0 && mol_nshells && shells, XCQ_ERROR_INVALID_ARGUMENT, "empty basis");
  int32_t nshells = 0;
  for (int32_t m = 0; m < nmol; ++m) {
    XCQ_REQUIRE(mol_nshells[m] > 0, XCQ_ERROR_INVALID_ARGUMENT, "molecule without shells");
    nshells += mol_nshells[m];
  }
  XCQ_HIP(hipSetDevice(ctx->device));
  std::vector<DevShell> h(nshells);
  std::vector<MolInfo> mols(nmol);
  ctx->sh_size.assign(nshells, 0);
  ctx->sh_first.assign(nshells, 0);
  ctx->sh_mol.assign(nshells, 0);
  ctx->sh_origin.assign(3 * (size_t)nshells, 0.);
  int32_t i = 0;
  int64_t p_total = 0;
  int32_t max_nbf = 0, max_l = 0;
  for (int32_t m = 0; m < nmol; ++m) {
    mols[m].sh_begin = i;
    mols[m].nshells = mol_nshells[m];
    int32_t nbf = 0;
    for (int32_t k = 0; k < mol_nshells[m]; ++k, ++i) {
      const auto& s = shells[i];
      XCQ_REQUIRE(s.l >= 0 && s.l <= kMaxL, XCQ_ERROR_INVALID_ARGUMENT,
                  "shell angular momentum out of range (0..4)");
      XCQ_REQUIRE(s.nprim > 0 && s.nprim <= 32, XCQ_ERROR_INVALID_ARGUMENT,
                  "shell primitive count out of range (1..32)");
      auto& d = h[i];
      std::memset(&d, 0, sizeof(d));
      for (int q = 0; q < s.nprim; ++q) {
        d.alpha[q] = s.alpha[q];
        d.coeff[q] = s.coeff[q];
      }
      for (int q = 0; q < 3; ++q) d.O[q] = ctx->sh_origin[3 * (size_t)i + q] = s.origin[q];
      d.nprim = s.nprim;
      d.l = s.l;
      d.pure = s.pure;
      d.size = s.pure ? 2 * s.l + 1 : (s.l + 1) * (s.l + 2) / 2;
      max_l = std::max(max_l, s.l);
      ctx->sh_size[i] = d.size;
      ctx->sh_first[i] = nbf;
      ctx->sh_mol[i] = m;
      nbf += d.size;
    }
    mols[m].nbf = nbf;
    mols[m].p_off = p_total;
    p_total += (int64_t)nbf * nbf;
    max_nbf = std::max(max_nbf, nbf);
  }
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  ctx->drop_graph();
  ctx->free_tasks();  // tasks refer to shell indices: require a new upload
  dfree(ctx->d_shells);
  dfree(ctx->d_shc);
  int rc = dalloc(&ctx->d_shells, nshells);
  if (rc) return rc;
  XCQ_HIP(hipMemcpy(ctx->d_shells, h.data(), sizeof(DevShell) * nshells, hipMemcpyHostToDevice));
  std::vector<DevShellC> hc(nshells);
  for (int32_t k = 0; k < nshells; ++k) {
    const DevShell& d = h[k];
    hc[k] = DevShellC{d.O[0], d.O[1], d.O[2], d.alpha[0], d.coeff[0],
                      d.l | (d.pure ? 1 << 8 : 0) | (d.nprim << 16), 0};
  }
  if ((rc = dalloc(&ctx->d_shc, nshells))) return rc;
  XCQ_HIP(hipMemcpy(ctx->d_shc, hc.data(), sizeof(DevShellC) * nshells, hipMemcpyHostToDevice));
  ctx->nshells = nshells;
  const bool same_shape =
      ctx->mols.size() == mols.size() &&
      std::equal(mols.begin(), mols.end(), ctx->mols.begin(),
                 [](const MolInfo& a, const MolInfo& b) { return a.nbf == b.nbf; });
  if (!same_shape) {
    dfree(ctx->d_P);
    dfree(ctx->d_Pf32);
    dfree(ctx->d_vxc);
    dfree(ctx->d_Pz);
    dfree(ctx->d_vxcz);
    ctx->have_pz = false;
    dfree(ctx->d_tP);
    ctx->have_tp = false;
    dfree(ctx->d_tPz);
    ctx->have_tpz = false;
    dfree(ctx->d_out2);
    dfree(ctx->d_corr);
    dfree(ctx->d_rpart);
    dfree(ctx->d_mol_n);
    dfree(ctx->d_mol_off);
    if ((rc = dalloc(&ctx->d_P, (size_t)p_total))) return rc;
    if ((rc = dalloc(&ctx->d_vxc, (size_t)p_total))) return rc;
    if ((rc = dalloc(&ctx->d_out2, (size_t)2 * nmol))) return rc;
    if ((rc = dalloc(&ctx->d_corr, (size_t)kCorrBlocks * nmol))) return rc;
    if ((rc = dalloc(&ctx->d_rpart, (size_t)2 * kRedBlocks * nmol))) return rc;
    if ((rc = dalloc(&ctx->d_mol_n, (size_t)nmol))) return rc;
    if ((rc = dalloc(&ctx->d_mol_off, (size_t)nmol))) return rc;
    std::vector<int32_t> mn(nmol);
    std::vector<int64_t> mo(nmol);
    for (int32_t m = 0; m < nmol; ++m) {
      mn[m] = mols[m].nbf;
      mo[m] = mols[m].p_off;
    }
    XCQ_HIP(hipMemcpy(ctx->d_mol_n, mn.data(), 4 * (size_t)nmol, hipMemcpyHostToDevice));
    XCQ_HIP(hipMemcpy(ctx->d_mol_off, mo.data(), 8 * (size_t)nmol, hipMemcpyHostToDevice));
  } else {
    for (int32_t m = 0; m < nmol; ++m) mols[m].have_density = ctx->mols[m].have_density;
  }
  ctx->mols = std::move(mols);
  ctx->p_total = p_total;
  ctx->max_nbf = max_nbf;
  ctx->max_l = max_l;
  return XCQ_SUCCESS;
}

int xcq_upload_basis(xcq_context ctx, int32_t nshells, const xcq_shell* shells) {
  XCQ_REQUIRE(nshells > 0 && shells, XCQ_ERROR_INVALID_ARGUMENT, "empty basis");
  return xcq_upload_molecules(ctx, 1, &nshells, shells);
}

int xcq_upload_tasks_batched(xcq_context ctx, int32_t ntasks, const xcq_task* tasks,
                             const int32_t* task_mol, int64_t npts_total, const double* points,
                             const double* weights, int64_t nshell_entries,
                             const int32_t* shell_lists) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(ctx->nshells > 0, XCQ_ERROR_STATE, "basis must be uploaded before tasks");
  XCQ_REQUIRE(ntasks >= 0 && (ntasks == 0 || tasks), XCQ_ERROR_INVALID_ARGUMENT, "bad task table");
  XCQ_REQUIRE(npts_total >= 0 && (npts_total == 0 || (points && weights)),
              XCQ_ERROR_INVALID_ARGUMENT, "bad point arrays");
  XCQ_REQUIRE(nshell_entries >= 0 && (nshell_entries == 0 || shell_lists),
              XCQ_ERROR_INVALID_ARGUMENT, "bad shell lists");
  XCQ_HIP(hipSetDevice(ctx->device));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  const int32_t nmol = (int32_t)ctx->mols.size();

  // Validate tasks and build the per-task device descriptors.
  std::vector<DevTask> ht(ntasks);
  std::vector<int32_t> bfmap, lsh;
  std::vector<int64_t> mol_blocks(nmol + 1, 0);
  for (int32_t i = 0; i < ntasks; ++i) {
    const auto& t = tasks[i];
    const int32_t m = task_mol ? task_mol[i] : 0;
    XCQ_REQUIRE(m >= 0 && m < nmol, XCQ_ERROR_INVALID_ARGUMENT, "task molecule out of range");
    XCQ_REQUIRE(t.npts >= 0 && t.nshells >= 0, XCQ_ERROR_INVALID_ARGUMENT, "negative task size");
    XCQ_REQUIRE(t.point_offset >= 0 && t.point_offset + t.npts <= npts_total,
                XCQ_ERROR_INVALID_ARGUMENT, "task points out of range");
    XCQ_REQUIRE(t.shell_offset >= 0 && t.shell_offset + t.nshells <= nshell_entries,
                XCQ_ERROR_INVALID_ARGUMENT, "task shell list out of range");
    auto& d = ht[i];
    std::memset(&d, 0, sizeof(d));
    d.npts = t.npts;
    d.nshells = t.nshells;
    d.iparent = t.iparent;
    d.npts_pad = (t.npts + kPtBlock - 1) / kPtBlock * kPtBlock;
    d.mol = m;
    d.ldp = ctx->mols[m].nbf;
    d.p_off = ctx->mols[m].p_off;
    d.pt_off = t.point_offset;
    d.sh_off = t.shell_offset;
    d.bf_off = (int64_t)bfmap.size();
    d.lsh_off = (int64_t)lsh.size();
    int32_t nbe = 0, prev = -1;
    for (int32_t k = 0; k < t.nshells; ++k) {
      const int32_t s = shell_lists[t.shell_offset + k];
      XCQ_REQUIRE(s >= 0 && s < ctx->nshells, XCQ_ERROR_INVALID_ARGUMENT,
                  "shell index out of range");
      XCQ_REQUIRE(ctx->sh_mol[s] == m, XCQ_ERROR_INVALID_ARGUMENT,
                  "task shell list names a shell of another molecule");
      XCQ_REQUIRE(s > prev, XCQ_ERROR_INVALID_ARGUMENT, "shell lists must be ascending");
      prev = s;
      lsh.push_back(nbe);
      for (int32_t f = 0; f < ctx->sh_size[s]; ++f) bfmap.push_back(ctx->sh_first[s] + f);
      nbe += ctx->sh_size[s];
    }
    XCQ_REQUIRE(t.nbe == nbe, XCQ_ERROR_INVALID_ARGUMENT,
                "task nbe does not match its shell list");
    d.nbe = nbe;
    mol_blocks[m + 1] += d.npts_pad / kPtBlock;
  }
  // Point-block partials are numbered molecule by molecule so that each
  // molecule's EXC/NEL is one contiguous reduction.
  for (int32_t m = 0; m < nmol; ++m) mol_blocks[m + 1] += mol_blocks[m];
  {
    std::vector<int64_t> next(mol_blocks.begin(), mol_blocks.end() - 1);
    for (auto& d : ht) {
      d.red_off = next[d.mol];
      next[d.mol] += d.npts_pad / kPtBlock;
    }
  }
  int64_t tmax_total = 0;
  for (auto& d : ht) {
    d.tmax_off = tmax_total;
    tmax_total += (d.nbe + kGran - 1) / kGran;
  }
  const int64_t pblocks = mol_blocks[nmol];

  // Chunk planning against the arena.
  size_t free_b = 0, total_b = 0;
  XCQ_HIP(hipMemGetInfo(&free_b, &total_b));
  const size_t resident = (size_t)npts_total * 32 + bfmap.size() * 4 + lsh.size() * 4 +
                          (size_t)nshell_entries * 4 + (size_t)ntasks * sizeof(DevTask) +
                          (size_t)pblocks * 16 * 8;
  size_t budget;
  if (ctx->user_arena) {
    budget = ctx->user_arena_bytes;
  } else if (ctx->arena_bytes) {
    budget = ctx->arena_bytes;
  } else {
    const double avail = ctx->fill * (double)free_b - (double)resident;
    budget = (size_t)std::max(0.0, std::min(avail * 0.5, 128.0 * (1ull << 30)));
  }
  // FP64 storage needs 40 bytes per (point, function) of a task, FP32 storage
  // 22 (five float planes + the value plane's bfloat16 residual); the arena
  // is sized for the larger.
  auto task_bytes = [](const DevTask& d) { return (size_t)5 * d.nbe * d.npts_pad * 8; };
  std::vector<Chunk> chunks;
  int64_t max_bytes = 0;
  {
    int32_t i = 0;
    while (i < ntasks) {
      Chunk c;
      c.task_begin = i;
      size_t used = 0;
      int64_t se = 0, se_mx = 0;
      while (i < ntasks) {
        auto& d = ht[i];
        const size_t b = task_bytes(d);
        // Kernels address a task's planes with 32-bit byte offsets.
        XCQ_REQUIRE(b < (size_t(1) << 32), XCQ_ERROR_INVALID_ARGUMENT,
                    "task too large: its basis function and point planes exceed 4 GiB");
        if (used + b > budget) {
          XCQ_REQUIRE(i > c.task_begin, XCQ_ERROR_OUT_OF_MEMORY,
                      "Insufficient memory to even start!");
          break;
        }
        used += b;
        const int64_t plane = (int64_t)d.nbe * d.npts_pad;  // npts_pad is a multiple of 64
        d.b_off = se;
        d.b_off_mx = se_mx;
        se += 5 * plane;
        se_mx += 5 * plane + plane / 2;
        ++i;
      }
      c.task_end = i;
      c.scratch_elems = se;
      max_bytes = std::max(max_bytes, std::max(se * 8, se_mx * 4));
      chunks.push_back(c);
    }
  }

  // Release the previous task set and allocate the new one.
  ctx->drop_graph();
  ctx->free_tasks();
  std::vector<double> hx(npts_total), hy(npts_total), hz(npts_total);
  for (int64_t k = 0; k < npts_total; ++k) {
    hx[k] = points[3 * k];
    hy[k] = points[3 * k + 1];
    hz[k] = points[3 * k + 2];
  }
  int rc;
  if ((rc = dalloc(&ctx->d_tasks, ht.size()))) return rc;
  if ((rc = dalloc(&ctx->d_px, npts_total))) return rc;
  if ((rc = dalloc(&ctx->d_py, npts_total))) return rc;
  if ((rc = dalloc(&ctx->d_pz, npts_total))) return rc;
  if ((rc = dalloc(&ctx->d_w, npts_total))) return rc;
  if ((rc = dalloc(&ctx->d_shell_list, (size_t)nshell_entries))) return rc;
  bfmap.resize(bfmap.size() + kBfPad, 0);
  if ((rc = dalloc(&ctx->d_bfmap, bfmap.size()))) return rc;
  if ((rc = dalloc(&ctx->d_lsh, lsh.size()))) return rc;
  if ((rc = dalloc(&ctx->d_red, (size_t)2 * std::max<int64_t>(pblocks, 1)))) return rc;
  ctx->pblocks = pblocks;
  dfree(ctx->d_tmax);
  if ((rc = dalloc(&ctx->d_tmax, (size_t)std::max<int64_t>(tmax_total, 1)))) return rc;
  ctx->tmax_total = tmax_total;
  dfree(ctx->d_mol_blk);
  if ((rc = dalloc(&ctx->d_mol_blk, mol_blocks.size()))) return rc;
  ctx->scratch_bytes = (size_t)max_bytes;
  if (ctx->user_arena) {
    // Every chunk fits the caller's buffer (the planner's budget).
    ctx->d_scratch = ctx->user_arena;
  } else if (ctx->scratch_bytes) {
    hipError_t e = hipMalloc(&ctx->d_scratch, ctx->scratch_bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      set_error("hipMalloc of the chunk scratch failed: " + std::string(hipGetErrorString(e)));
      return XCQ_ERROR_OUT_OF_MEMORY;
    }
  }
  auto h2d = [](void* d, const void* h, size_t b) {
    return b ? hipMemcpy(d, h, b, hipMemcpyHostToDevice) : hipSuccess;
  };
  XCQ_HIP(h2d(ctx->d_px, hx.data(), npts_total * 8));
  XCQ_HIP(h2d(ctx->d_py, hy.data(), npts_total * 8));
  XCQ_HIP(h2d(ctx->d_pz, hz.data(), npts_total * 8));
  XCQ_HIP(h2d(ctx->d_w, weights, npts_total * 8));
  XCQ_HIP(h2d(ctx->d_shell_list, shell_lists, nshell_entries * 4));
  XCQ_HIP(h2d(ctx->d_bfmap, bfmap.data(), bfmap.size() * 4));
  XCQ_HIP(h2d(ctx->d_lsh, lsh.data(), lsh.size() * 4));
  XCQ_HIP(h2d(ctx->d_mol_blk, mol_blocks.data(), mol_blocks.size() * 8));

  ctx->ev.resize(chunks.size() * kStages * 2);
  for (auto& e : ctx->ev) XCQ_HIP(hipEventCreate(&e));

  ctx->h_tasks = std::move(ht);
  ctx->chunks = std::move(chunks);
  ctx->ntasks = ntasks;
  ctx->npts_total = npts_total;
  ctx->total_pblocks = pblocks;
  ctx->cost_d.clear();
  ctx->cost_s.clear();
  if ((rc = build_work_lists(ctx))) return rc;
  XCQ_HIP(h2d(ctx->d_tasks, ctx->h_tasks.data(), ctx->h_tasks.size() * sizeof(DevTask)));
  int64_t sum_pts = 0, sum_pnn = 0;
  for (const auto& d : ctx->h_tasks) {
    sum_pts += d.npts;
    sum_pnn += (int64_t)d.npts * d.nbe * d.nbe;
  }
  ctx->counters[2] = sum_pts;
  ctx->counters[3] = sum_pnn;
  return XCQ_SUCCESS;
}

int xcq_upload_tasks(xcq_context ctx, int32_t ntasks, const xcq_task* tasks, int64_t npts_total,
                     const double* points, const double* weights, int64_t nshell_entries,
                     const int32_t* shell_lists) {
  return xcq_upload_tasks_batched(ctx, ntasks, tasks, nullptr, npts_total, points, weights,
                                  nshell_entries, shell_lists);
}

int xcq_set_tile_policy(xcq_context ctx, int32_t dens32_max_nbe, int32_t syr2k32_max_nbe) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_HIP(hipSetDevice(ctx->device));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  ctx->dens32_max_nbe = dens32_max_nbe;
  ctx->syr2k32_max_nbe = syr2k32_max_nbe;
  if (ctx->ntasks == 0) return XCQ_SUCCESS;
  int rc = build_work_lists(ctx);
  if (rc) return rc;
  XCQ_HIP(hipMemcpy(ctx->d_tasks, ctx->h_tasks.data(), ctx->h_tasks.size() * sizeof(DevTask),
                    hipMemcpyHostToDevice));
  return XCQ_SUCCESS;
}

int xcq_get_tile_policy(xcq_context ctx, int32_t* out5) {
  XCQ_REQUIRE(ctx && out5, XCQ_ERROR_INVALID_ARGUMENT, "null argument");
  out5[0] = ctx->dens32_max_nbe;
  out5[1] = ctx->syr2k32_max_nbe;
  out5[2] = ctx->dens_hiocc_max_nbe;
  out5[3] = ctx->nqueues;
  out5[4] = ctx->dens128_min_nbe;
  return XCQ_SUCCESS;
}

int xcq_set_ssf_neighbour_lists(xcq_context ctx, int32_t enable) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  ctx->ssf_lists = enable != 0;
  return XCQ_SUCCESS;
}

int xcq_set_vxc_drop_tolerance(xcq_context ctx, double tol) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(tol >= 0.0 && tol <= 1e-10, XCQ_ERROR_INVALID_ARGUMENT,
              "VXC drop tolerance must lie in [0, 1e-10]");
  XCQ_HIP(hipSetDevice(ctx->device));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  ctx->vxc_drop = tol;
  ctx->drop_graph();  // the captured launches carry the old value
  return XCQ_SUCCESS;
}

int xcq_set_density_wide(xcq_context ctx, int32_t wide_min_nbe) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_HIP(hipSetDevice(ctx->device));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  ctx->dens128_min_nbe = wide_min_nbe;
  if (ctx->ntasks == 0) return XCQ_SUCCESS;
  int rc = build_work_lists(ctx);
  if (rc) return rc;
  XCQ_HIP(hipMemcpy(ctx->d_tasks, ctx->h_tasks.data(), ctx->h_tasks.size() * sizeof(DevTask),
                    hipMemcpyHostToDevice));
  return XCQ_SUCCESS;
}

int xcq_set_density_basis_eval(xcq_context ctx, int32_t in_stage) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_HIP(hipSetDevice(ctx->device));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  XCQ_REQUIRE(in_stage == 0 || in_stage == 1 || in_stage == 3, XCQ_ERROR_INVALID_ARGUMENT,
              "density basis evaluation mode must be 0, 1 or 3");
  ctx->dens_in_stage = in_stage;
  ctx->drop_graph();  // the captured launches follow the old setting
  return XCQ_SUCCESS;
}

int xcq_get_density_basis_eval(xcq_context ctx, int32_t* mode) {
  XCQ_REQUIRE(ctx && mode, XCQ_ERROR_INVALID_ARGUMENT, "null argument");
  *mode = ctx->max_l <= kInStageMaxL ? ctx->dens_in_stage : 0;
  return XCQ_SUCCESS;
}

int xcq_set_density_occupancy(xcq_context ctx, int32_t hiocc_max_nbe) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_HIP(hipSetDevice(ctx->device));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  ctx->dens_hiocc_max_nbe = hiocc_max_nbe;
  if (ctx->ntasks == 0) return XCQ_SUCCESS;
  int rc = build_work_lists(ctx);
  if (rc) return rc;
  XCQ_HIP(hipMemcpy(ctx->d_tasks, ctx->h_tasks.data(), ctx->h_tasks.size() * sizeof(DevTask),
                    hipMemcpyHostToDevice));
  return XCQ_SUCCESS;
}

int xcq_set_schedule(xcq_context ctx, int32_t nqueues, const double* dens_cost,
                     const double* syr2k_cost) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(nqueues >= 0, XCQ_ERROR_INVALID_ARGUMENT, "negative queue count");
  XCQ_REQUIRE((dens_cost == nullptr) == (syr2k_cost == nullptr), XCQ_ERROR_INVALID_ARGUMENT,
              "give both cost arrays or neither");
  XCQ_HIP(hipSetDevice(ctx->device));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  ctx->nqueues = nqueues;
  if (dens_cost) {
    XCQ_REQUIRE(ctx->ntasks > 0, XCQ_ERROR_STATE, "costs need an uploaded task set");
    ctx->cost_d.assign(dens_cost, dens_cost + ctx->ntasks);
    ctx->cost_s.assign(syr2k_cost, syr2k_cost + ctx->ntasks);
  } else {
    ctx->cost_d.clear();
    ctx->cost_s.clear();
  }
  if (ctx->ntasks == 0) return XCQ_SUCCESS;
  return build_work_lists(ctx);
}

int xcq_enable_task_timing(xcq_context ctx, int enable) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_HIP(hipSetDevice(ctx->device));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  ctx->timing = enable != 0;
  dfree(ctx->d_tstamp);
  ctx->drop_graph();
  if (ctx->timing && !ctx->h_work.empty()) {
    int rc = dalloc(&ctx->d_tstamp, 2 * ctx->h_work.size());
    if (rc) return rc;
    XCQ_HIP(hipMemset(ctx->d_tstamp, 0, 2 * ctx->h_work.size() * sizeof(int64_t)));
  }
  return XCQ_SUCCESS;
}

int xcq_task_times(xcq_context ctx, double* dens_us, double* syr2k_us) {
  XCQ_REQUIRE(ctx && dens_us && syr2k_us, XCQ_ERROR_INVALID_ARGUMENT, "null argument");
  XCQ_REQUIRE(ctx->timing && ctx->d_tstamp, XCQ_ERROR_STATE, "task timing is not enabled");
  XCQ_HIP(hipSetDevice(ctx->device));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<int64_t> ts(2 * ctx->h_work.size());
  XCQ_HIP(hipMemcpy(ts.data(), ctx->d_tstamp, ts.size() * 8, hipMemcpyDeviceToHost));
  std::fill(dens_us, dens_us + ctx->ntasks, 0.0);
  std::fill(syr2k_us, syr2k_us + ctx->ntasks, 0.0);
  const double us_per_tick = 1000.0 / ctx->wall_khz;
  auto add = [&](const WorkRange& r, double* out) {
    for (int64_t k = r.off; k < r.off + r.n; ++k) {
      const int64_t dt = ts[2 * k + 1] - ts[2 * k];
      if (dt > 0) out[ctx->h_work[k].task] += dt * us_per_tick;
    }
  };
  for (const auto& c : ctx->chunks) {
    for (int v = 0; v < 4; ++v) add(c.dens[v], dens_us);
    for (int v = 0; v < 2; ++v) add(c.syr2k[v], syr2k_us);
  }
  return XCQ_SUCCESS;
}

int xcq_use_graph(xcq_context ctx, int enable) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_HIP(hipSetDevice(ctx->device));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  ctx->use_graph = enable != 0;
  if (!ctx->use_graph) ctx->drop_graph();
  return XCQ_SUCCESS;
}

int xcq_autotune(xcq_context ctx, int functional, int precision, int32_t* dens32_max_nbe,
                 int32_t* syr2k32_max_nbe, int32_t* nqueues, double* best_ms) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(ctx->ntasks > 0, XCQ_ERROR_STATE, "no tasks uploaded");
  const bool graph = ctx->use_graph;
  ctx->use_graph = false;  // stage timing needs the eager path
  ctx->drop_graph();
  // Median of three timed runs of (density, B^T Z) stage times.
  auto measure = [&](double* dens_ms, double* syr_ms, double* tot_ms) -> int {
    std::vector<double> d, sy, tt;
    for (int rep = 0; rep < 4; ++rep) {
      int rc = run_pipeline(ctx, functional, precision, 1);
      if (rc) return rc;
      double st[kStages + 1];
      xcq_stage_times(ctx, st, kStages + 1);
      if (rep == 0) continue;  // warm-up
      d.push_back(st[0] + st[1] + st[3]);
      sy.push_back(st[4]);
      tt.push_back(st[kStages]);
    }
    auto med = [](std::vector<double>& v) {
      std::sort(v.begin(), v.end());
      return v[v.size() / 2];
    };
    *dens_ms = med(d);
    *syr_ms = med(sy);
    *tot_ms = med(tt);
    return XCQ_SUCCESS;
  };
  const int32_t cand[] = {0, 32, 48, 64, 96, 128, 192, 256, 1 << 30};
  int32_t best_d = ctx->dens32_max_nbe, best_s = ctx->syr2k32_max_nbe;
  double bd = 1e300, bs = 1e300, tot = 0.;
  int rc;
  // The density-side times below are collocation + density + the Z launch:
  // the basis evaluation mode moves work between them.  The two stages' tile
  // thresholds are disjoint: sweep both at once and keep the best of each.
  for (int32_t c : cand) {
    if ((rc = xcq_set_tile_policy(ctx, c, c))) return rc;
    double dm, sm;
    if ((rc = measure(&dm, &sm, &tot))) return rc;
    if (dm < bd) {
      bd = dm;
      best_d = c;
    }
    if (sm < bs) {
      bs = sm;
      best_s = c;
    }
  }
  if ((rc = xcq_set_tile_policy(ctx, best_d, best_s))) return rc;
  {
    // Per basis evaluation mode (in-stage first, when the basis allows it):
    // the 32-wide threshold, which 64-wide tasks run at three workgroups per
    // CU (mixed precision, and the FP64 in-stage kernel) and, for mixed
    // precision, from which nbe on 128-wide tiles are used; the mode with
    // the best collocation + density (+ Z launch) time wins.
    const bool mixed = precision == XCQ_PREC_MIXED;
    struct Pol {
      int32_t mode, d, h, w;
      double ms;
    };
    std::vector<int32_t> modes;
    if (ctx->max_l <= kInStageMaxL) modes.push_back(3);
    modes.push_back(0);
    Pol best{ctx->dens_in_stage, best_d, ctx->dens_hiocc_max_nbe, ctx->dens128_min_nbe, 1e300};
    for (int32_t mode : modes) {
      if ((rc = xcq_set_density_basis_eval(ctx, mode))) return rc;
      Pol p{mode, best_d, 0, 1 << 30, 1e300};
      if ((rc = xcq_set_density_occupancy(ctx, 0))) return rc;
      if ((rc = xcq_set_density_wide(ctx, 1 << 30))) return rc;
      for (int32_t c : cand) {
        if ((rc = xcq_set_tile_policy(ctx, c, best_s))) return rc;
        double dm, sm;
        if ((rc = measure(&dm, &sm, &tot))) return rc;
        if (dm < p.ms) {
          p.ms = dm;
          p.d = c;
        }
      }
      if ((rc = xcq_set_tile_policy(ctx, p.d, best_s))) return rc;
      if ((mixed && mode != 1) || (!mixed && mode == 3)) {
        for (int32_t c : cand) {
          if ((rc = xcq_set_density_occupancy(ctx, c))) return rc;
          double dm, sm;
          if ((rc = measure(&dm, &sm, &tot))) return rc;
          if (dm < p.ms) {
            p.ms = dm;
            p.h = c;
          }
        }
        if ((rc = xcq_set_density_occupancy(ctx, p.h))) return rc;
      }
      for (int32_t c : {128, 192, 256, 384, 512, 1 << 30}) {
        if (!mixed) break;
        if ((rc = xcq_set_density_wide(ctx, c))) return rc;
        double dm, sm;
        if ((rc = measure(&dm, &sm, &tot))) return rc;
        if (dm < p.ms) {
          p.ms = dm;
          p.w = c;
        }
      }
      if (p.ms < best.ms) best = p;
    }
    // A second pass over the 32-wide and the occupancy thresholds in the
    // winning mode, the other settings at their chosen values: the first
    // pass picked the 32-wide threshold with every 64-wide task at two
    // workgroups per CU, and the two thresholds interact.
    if ((rc = xcq_set_density_basis_eval(ctx, best.mode))) return rc;
    if ((rc = xcq_set_density_occupancy(ctx, best.h))) return rc;
    if ((rc = xcq_set_density_wide(ctx, best.w))) return rc;
    for (int32_t c : cand) {
      if ((rc = xcq_set_tile_policy(ctx, c, best_s))) return rc;
      double dm, sm;
      if ((rc = measure(&dm, &sm, &tot))) return rc;
      if (dm < best.ms) {
        best.ms = dm;
        best.d = c;
      }
    }
    if ((rc = xcq_set_tile_policy(ctx, best.d, best_s))) return rc;
    if ((mixed && best.mode != 1) || (!mixed && best.mode == 3)) {
      for (int32_t c : cand) {
        if ((rc = xcq_set_density_occupancy(ctx, c))) return rc;
        double dm, sm;
        if ((rc = measure(&dm, &sm, &tot))) return rc;
        if (dm < best.ms) {
          best.ms = dm;
          best.h = c;
        }
      }
    }
    best_d = best.d;
    if ((rc = xcq_set_tile_policy(ctx, best.d, best_s))) return rc;
    if ((rc = xcq_set_density_occupancy(ctx, best.h))) return rc;
  }
  // Persistent workgroups on dynamic queues (at least as many as any stage
  // kernel keeps resident per CU; surplus ones find the queues empty)
  // against one workgroup per item.
  int dev_cus = 0;
  if (hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
      hipSuccess) {
    (void)hipGetLastError();
    dev_cus = 256;
  }
  const std::vector<double> keep_d = ctx->cost_d, keep_s = ctx->cost_s;
  int32_t best_q = 0;
  double bt = 1e300;
  for (int32_t q : {0, 4 * dev_cus, 8 * dev_cus}) {
    if ((rc = xcq_set_schedule(ctx, q, keep_d.empty() ? nullptr : keep_d.data(),
                               keep_s.empty() ? nullptr : keep_s.data())))
      return rc;
    double dm, sm, tm;
    if ((rc = measure(&dm, &sm, &tm))) return rc;
    if (tm < bt) {
      bt = tm;
      best_q = q;
    }
  }
  if ((rc = xcq_set_schedule(ctx, best_q, keep_d.empty() ? nullptr : keep_d.data(),
                             keep_s.empty() ? nullptr : keep_s.data())))
    return rc;
  ctx->use_graph = graph;
  if (dens32_max_nbe) *dens32_max_nbe = best_d;
  if (syr2k32_max_nbe) *syr2k32_max_nbe = best_s;
  if (nqueues) *nqueues = best_q;
  if (best_ms) *best_ms = bt;
  return XCQ_SUCCESS;
}

int xcq_upload_density_mol(xcq_context ctx, int32_t mol, int32_t nbf, const double* P,
                           int64_t ldp) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(P, XCQ_ERROR_INVALID_ARGUMENT, "null density");
  XCQ_REQUIRE(mol >= 0 && mol < (int32_t)ctx->mols.size(), XCQ_ERROR_INVALID_ARGUMENT,
              "molecule index out of range");
  MolInfo& M = ctx->mols[mol];
  XCQ_REQUIRE(nbf == M.nbf && nbf > 0, XCQ_ERROR_INVALID_ARGUMENT,
              "density dimension does not match the basis");
  XCQ_REQUIRE(ldp >= nbf, XCQ_ERROR_INVALID_ARGUMENT, "ldp < nbf");
  XCQ_HIP(hipSetDevice(ctx->device));
  XCQ_HIP(hipMemcpy2DAsync(ctx->d_P + M.p_off, (size_t)nbf * 8, P, (size_t)ldp * 8,
                           (size_t)nbf * 8, nbf, hipMemcpyHostToDevice, ctx->stream));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  M.have_density = true;
  ctx->pf32_valid = false;  // the FP32 copy follows the new density
  return XCQ_SUCCESS;
}

int xcq_upload_density(xcq_context ctx, int32_t nbf, const double* P, int64_t ldp) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(ctx->mols.size() == 1, XCQ_ERROR_STATE,
              "context holds several molecules: use xcq_upload_density_mol");
  return xcq_upload_density_mol(ctx, 0, nbf, P, ldp);
}

int xcq_exc_vxc(xcq_context ctx, int functional, int precision, int want_vxc) {
  XCQ_REQUIRE(functional != 0, XCQ_ERROR_INVALID_ARGUMENT, "unknown functional");
  return run_pipeline(ctx, functional, precision, want_vxc ? 1 : 0);
}

// ---- UKS (GauXC's eval_exc_vxc(Ps, Pz)) --------------------------------------
namespace {

// Largest molecule whose neighbour lists (12 bytes per atom pair) the SSF
// launches build; above it they loop over every atom.
constexpr int32_t kSsfNbrMaxAtoms = 8192;

// Per atom A, every atom in order of increasing R_AB with A itself first
// (row-major natoms x natoms), for k_ssf's bounded loops.
void ssf_neighbours(int32_t natoms, const std::vector<double>& rab, std::vector<int32_t>& idx,
                    std::vector<double>& r) {
  idx.resize((size_t)natoms * natoms);
  r.resize(idx.size());
  std::vector<int32_t> ord(natoms);
  for (int32_t A = 0; A < natoms; ++A) {
    const double* col = rab.data() + (size_t)A * natoms;  // R_AB = rab[B + A * natoms]
    std::iota(ord.begin(), ord.end(), 0);
    std::swap(ord[0], ord[A]);
    std::stable_sort(ord.begin() + 1, ord.end(),
                     [&](int32_t u, int32_t v) { return col[u] < col[v]; });
    for (int32_t k = 0; k < natoms; ++k) {
      idx[(size_t)A * natoms + k] = ord[k];
      r[(size_t)A * natoms + k] = k == 0 ? 0.0 : col[ord[k]];
    }
  }
}

// R_AB (column-major) and nearest-neighbour distances of MolMeta.
void ssf_molmeta(int32_t natoms, const double* atom_xyz, std::vector<double>& rab,
                 std::vector<double>& dn) {
  rab.assign((size_t)natoms * natoms, 0.);
  dn.assign(natoms, INFINITY);
  for (int i = 0; i < natoms; ++i)
    for (int j = 0; j < natoms; ++j) {
      const double dx = atom_xyz[3 * i] - atom_xyz[3 * j];
      const double dy = atom_xyz[3 * i + 1] - atom_xyz[3 * j + 1];
      const double dz = atom_xyz[3 * i + 2] - atom_xyz[3 * j + 2];
      rab[(size_t)i + (size_t)j * natoms] = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (i != j) dn[i] = std::min(dn[i], rab[(size_t)i + (size_t)j * natoms]);
    }
}


// The per-point buffers of the UKS, FXC and gradient paths (freed with the
// task set, whose point-block count sizes them).
int ensure_point_buffers(xcq_context_s* ctx) {
  if (ctx->d_uks) return XCQ_SUCCESS;
  return dalloc(&ctx->d_uks, (size_t)5 * 256 * std::max<int64_t>(ctx->pblocks, 1));
}

// Basis-evaluation mode of the FP64 UKS and FXC paths: in-stage (3: the
// density stage evaluates B and grad B itself and Z is formed by
// k_z_in_stage from the per-point coefficients, so collocation stores the
// value plane only) up to kInStageMaxL, stored planes (0) above.
int derivative_mode(const xcq_context_s* ctx) { return ctx->max_l <= kInStageMaxL ? 3 : 0; }

// Z of a chunk from per-point coefficients ([block][4][64]) into plane 4.
int z_from_point_coeffs(xcq_context_s* ctx, const CtxView& v, const Chunk& c, int mode, int gga,
                        const double* coef) {
  if (mode == 0) return launch_z_from_coeffs(v, c.coll, gga, coef);
  return launch_z_in_stage(ctx->stream, v.work + c.coll.off, c.coll.n, v.tasks, v.shells,
                           v.shell_list, v.lsh, v.px, v.py, v.pz, coef, v.scratch,
                           XCQ_PREC_FP64, nullptr, gga, ctx->max_l);
}

// Z from per-point coefficients, then its B^T Z + Z^T B scattered into
// v.vxc.  In the in-stage mode with a drop tolerance the Z launch also
// records the granule maxima of |B| and |Z| (the bound table, cleared
// first), and the contraction skips negligible tile pairs and drops
// negligible elements as the energy path does; on stored planes only the
// element drop applies.
int z_and_vxc(xcq_context_s* ctx, CtxView v, const Chunk& c, int mode, int gga,
              const double* coef) {
  int rc;
  unsigned* tm = reinterpret_cast<unsigned*>(ctx->d_tmax);
  v.vxc_drop = ctx->vxc_drop;
  if (mode == 3 && v.vxc_drop > 0.0 && tm) {
    if ((rc = launch_fill(ctx->stream, ctx->d_tmax, ctx->tmax_total)) ||
        (rc = launch_z_in_stage(ctx->stream, v.work + c.coll.off, c.coll.n, v.tasks, v.shells,
                                v.shell_list, v.lsh, v.px, v.py, v.pz, coef, v.scratch,
                                XCQ_PREC_FP64, tm, gga, ctx->max_l)))
      return rc;
    v.tmax = tm;
  } else {
    if ((rc = z_from_point_coeffs(ctx, v, c, mode, gga, coef))) return rc;
    v.tmax = nullptr;
  }
  for (int k = 0; k < 2; ++k)
    if ((rc = launch_syr2k(v, c.syr2k[k], k == 0 ? 32 : 64, XCQ_PREC_FP64))) return rc;
  return XCQ_SUCCESS;
}

// One UKS evaluation on the context's stream (FP64; in-stage basis
// evaluation up to f functions, stored planes above): per chunk,
// collocation; the density stage twice (P_s, then P_z; factor 1) leaving
// raw per-point sums; the polarised functional and Z coefficients per
// point; Z_s into the chunk's Z plane and its B^T Z + Z^T B into VXC_s; then
// Z_z into VXC_z, both with the energy path's drop screening (z_and_vxc).
int enqueue_uks(xcq_context_s* ctx, int func, int want_vxc) {
  const int gga = (func == XCQ_FUNC_PBE || func == XCQ_FUNC_PBE0) ? 1 : 0;
  const int mode = derivative_mode(ctx);
  CtxView v = view_of(ctx);
  v.in_stage = mode;
  v.tmax = nullptr;
  v.vxc_drop = 0.0;
  hipStream_t s = ctx->stream;
  const size_t blk = (size_t)256 * ctx->pblocks;
  double* ds = ctx->d_uks;
  double* dz = ds + blk;
  double* zs = dz + blk;
  double* zz = zs + blk;
  int rc;
  if (want_vxc && ((rc = launch_fill(s, ctx->d_vxc, ctx->p_total)) ||
                   (rc = launch_fill(s, ctx->d_vxcz, ctx->p_total))))
    return rc;
  if (ctx->ncounters &&
      (rc = launch_fill(s, reinterpret_cast<double*>(ctx->d_counters), ctx->ncounters)))
    return rc;
  for (const Chunk& c : ctx->chunks) {
    const WorkItem* cw = v.work + c.coll.off;
    if ((rc = launch_collocation_fp64(s, cw, c.coll.n, v.tasks, v.shells, v.shell_list, v.lsh,
                                      v.px, v.py, v.pz, static_cast<double*>(v.scratch),
                                      gga && mode == 0, ctx->max_l)))
      return rc;
    for (int pass = 0; pass < 2; ++pass)
      for (int k = 0; k < 4; ++k)
        if ((rc = launch_density(v, c.dens[k], k == 0 ? 32 : 64, false, XCQ_PREC_FP64, func, gga,
                                 1.0, 0, pass == 0 ? ctx->d_P : ctx->d_Pz,
                                 pass == 0 ? ds : dz)))
          return rc;
    if ((rc = launch_uks_points(v, c.coll, func, gga, ds, dz, zs, zz))) return rc;
    if (!want_vxc) continue;
    for (int comp = 0; comp < 2; ++comp) {
      CtxView vc = v;
      vc.vxc = comp == 0 ? ctx->d_vxc : ctx->d_vxcz;
      if ((rc = z_and_vxc(ctx, vc, c, mode, gga, comp == 0 ? zs : zz))) return rc;
    }
  }
  if (want_vxc) {
    for (double* m : {ctx->d_vxc, ctx->d_vxcz})
      if ((rc = launch_symmetrize(s, 1, ctx->max_nbf, ctx->d_mol_n, ctx->d_mol_off, m)))
        return rc;
  }
  return launch_reduce_sums(s, 1, ctx->d_mol_blk, ctx->d_red, ctx->d_rpart, nullptr, 1.0,
                            ctx->d_out2);
}

}  // namespace

int xcq_upload_density_z(xcq_context ctx, int32_t nbf, const double* Pz, int64_t ldp) {
  XCQ_REQUIRE(ctx && Pz, XCQ_ERROR_INVALID_ARGUMENT, "null argument");
  XCQ_REQUIRE(ctx->mols.size() == 1, XCQ_ERROR_STATE, "UKS needs a single-molecule context");
  const MolInfo& M = ctx->mols[0];
  XCQ_REQUIRE(nbf == M.nbf && nbf > 0, XCQ_ERROR_INVALID_ARGUMENT,
              "density dimension does not match the basis");
  XCQ_REQUIRE(ldp >= nbf, XCQ_ERROR_INVALID_ARGUMENT, "ldp < nbf");
  XCQ_HIP(hipSetDevice(ctx->device));
  int rc;
  if (!ctx->d_Pz && (rc = dalloc(&ctx->d_Pz, (size_t)ctx->p_total))) return rc;
  XCQ_HIP(hipMemcpy2DAsync(ctx->d_Pz, (size_t)nbf * 8, Pz, (size_t)ldp * 8, (size_t)nbf * 8, nbf,
                           hipMemcpyHostToDevice, ctx->stream));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  ctx->have_pz = true;
  return XCQ_SUCCESS;
}

int xcq_exc_vxc_uks(xcq_context ctx, int functional, int want_vxc) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(functional == XCQ_FUNC_SVWN5 || functional == XCQ_FUNC_PBE ||
                  functional == XCQ_FUNC_PBE0,
              XCQ_ERROR_INVALID_ARGUMENT, "unknown functional");
  XCQ_REQUIRE(ctx->ntasks > 0, XCQ_ERROR_STATE, "no tasks uploaded");
  XCQ_REQUIRE(ctx->mols.size() == 1, XCQ_ERROR_STATE, "UKS needs a single-molecule context");
  XCQ_REQUIRE(ctx->mols[0].have_density && ctx->have_pz, XCQ_ERROR_STATE,
              "UKS needs P_s (xcq_upload_density) and P_z (xcq_upload_density_z)");
  XCQ_HIP(hipSetDevice(ctx->device));
  int rc;
  if ((rc = ensure_point_buffers(ctx))) return rc;
  if (!ctx->d_vxcz && (rc = dalloc(&ctx->d_vxcz, (size_t)ctx->p_total))) return rc;
  ctx->last_want_vxc = want_vxc ? 1 : 0;
  ctx->last_vxc_eval = want_vxc ? 1 : 0;
  ctx->last_graph = false;
  ctx->last_uks = true;
  ctx->last_fxc = false;
  ctx->last_fxc_uks = false;
  ctx->last_grad = false;
  return enqueue_uks(ctx, functional, want_vxc ? 1 : 0);
}

int xcq_download_results_uks(xcq_context ctx, double* exc, double* nel, double* vxc_s,
                             double* vxc_z, int64_t ldvxc) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(ctx->last_uks, XCQ_ERROR_STATE, "the last evaluation was not a UKS one");
  XCQ_REQUIRE(!vxc_z || vxc_s, XCQ_ERROR_INVALID_ARGUMENT, "vxc_z needs vxc_s");
  int rc = xcq_download_results_mol(ctx, 0, exc, nel, vxc_s, ldvxc);
  if (rc || !vxc_z) return rc;
  const MolInfo& M = ctx->mols[0];
  XCQ_HIP(hipMemcpy2DAsync(vxc_z, (size_t)ldvxc * 8, ctx->d_vxcz, (size_t)M.nbf * 8,
                           (size_t)M.nbf * 8, M.nbf, hipMemcpyDeviceToHost, ctx->stream));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  return XCQ_SUCCESS;
}

// ---- FXC contraction (GauXC's eval_fxc_contraction(P, tP)) ------------------
namespace {

// One RKS FXC contraction on the context's stream (FP64; in-stage basis
// evaluation up to f functions, stored planes above): per chunk,
// collocation; the density stage twice (P, then the trial density; factor
// 2) leaving raw per-point sums; the second-derivative kernel and Z
// coefficients per point; Z into the chunk's Z plane and its B^T Z + Z^T B
// into d_vxc with the energy path's drop screening (z_and_vxc);
// symmetrised.
int enqueue_fxc(xcq_context_s* ctx, int func) {
  const int gga = (func == XCQ_FUNC_PBE || func == XCQ_FUNC_PBE0) ? 1 : 0;
  const int mode = derivative_mode(ctx);
  CtxView v = view_of(ctx);
  v.in_stage = mode;
  v.tmax = nullptr;
  v.vxc_drop = 0.0;
  hipStream_t s = ctx->stream;
  const size_t blk = (size_t)256 * ctx->pblocks;
  double* dp = ctx->d_uks;
  double* dt = dp + blk;
  double* zc = dt + blk;
  int rc;
  if ((rc = launch_fill(s, ctx->d_vxc, ctx->p_total))) return rc;
  if (ctx->ncounters &&
      (rc = launch_fill(s, reinterpret_cast<double*>(ctx->d_counters), ctx->ncounters)))
    return rc;
  for (const Chunk& c : ctx->chunks) {
    const WorkItem* cw = v.work + c.coll.off;
    if ((rc = launch_collocation_fp64(s, cw, c.coll.n, v.tasks, v.shells, v.shell_list, v.lsh,
                                      v.px, v.py, v.pz, static_cast<double*>(v.scratch),
                                      gga && mode == 0, ctx->max_l)))
      return rc;
    for (int pass = 0; pass < 2; ++pass)
      for (int k = 0; k < 4; ++k)
        if ((rc = launch_density(v, c.dens[k], k == 0 ? 32 : 64, false, XCQ_PREC_FP64, func, gga,
                                 2.0, 0, pass == 0 ? ctx->d_P : ctx->d_tP,
                                 pass == 0 ? dp : dt)))
          return rc;
    if ((rc = launch_fxc_points(v, c.coll, func, gga, dp, dt, zc))) return rc;
    if ((rc = z_and_vxc(ctx, v, c, mode, gga, zc))) return rc;
  }
  if ((rc = launch_symmetrize(s, 1, ctx->max_nbf, ctx->d_mol_n, ctx->d_mol_off, ctx->d_vxc)))
    return rc;
  return launch_reduce_sums(s, 1, ctx->d_mol_blk, ctx->d_red, ctx->d_rpart, nullptr, 1.0,
                            ctx->d_out2);
}

}  // namespace

int xcq_upload_trial_density(xcq_context ctx, int32_t nbf, const double* tP, int64_t ldp) {
  XCQ_REQUIRE(ctx && tP, XCQ_ERROR_INVALID_ARGUMENT, "null argument");
  XCQ_REQUIRE(ctx->mols.size() == 1, XCQ_ERROR_STATE,
              "the FXC contraction needs a single-molecule context");
  const MolInfo& M = ctx->mols[0];
  XCQ_REQUIRE(nbf == M.nbf && nbf > 0, XCQ_ERROR_INVALID_ARGUMENT,
              "density dimension does not match the basis");
  XCQ_REQUIRE(ldp >= nbf, XCQ_ERROR_INVALID_ARGUMENT, "ldp < nbf");
  XCQ_HIP(hipSetDevice(ctx->device));
  int rc;
  if (!ctx->d_tP && (rc = dalloc(&ctx->d_tP, (size_t)ctx->p_total))) return rc;
  XCQ_HIP(hipMemcpy2DAsync(ctx->d_tP, (size_t)nbf * 8, tP, (size_t)ldp * 8, (size_t)nbf * 8, nbf,
                           hipMemcpyHostToDevice, ctx->stream));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  ctx->have_tp = true;
  return XCQ_SUCCESS;
}

int xcq_fxc_contraction(xcq_context ctx, int functional) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(functional == XCQ_FUNC_SVWN5 || functional == XCQ_FUNC_PBE ||
                  functional == XCQ_FUNC_PBE0,
              XCQ_ERROR_INVALID_ARGUMENT, "unknown functional");
  XCQ_REQUIRE(ctx->ntasks > 0, XCQ_ERROR_STATE, "no tasks uploaded");
  XCQ_REQUIRE(ctx->mols.size() == 1, XCQ_ERROR_STATE,
              "the FXC contraction needs a single-molecule context");
  XCQ_REQUIRE(ctx->mols[0].have_density && ctx->have_tp, XCQ_ERROR_STATE,
              "the FXC contraction needs P (xcq_upload_density) and the trial density "
              "(xcq_upload_trial_density)");
  XCQ_HIP(hipSetDevice(ctx->device));
  int rc;
  if ((rc = ensure_point_buffers(ctx))) return rc;
  ctx->last_want_vxc = 0;
  ctx->last_vxc_eval = 0;
  ctx->last_graph = false;
  ctx->last_uks = false;
  ctx->last_fxc = true;
  ctx->last_fxc_uks = false;
  ctx->last_grad = false;
  return enqueue_fxc(ctx, functional);
}

namespace {

// One UKS FXC contraction on the context's stream (FP64; basis mode as
// the RKS one): per chunk, collocation; the density stage four times (P_s,
// P_z, tP_s, tP_z; factor 1) leaving raw per-point sums; the derivatives of
// the UKS Z coefficients along the trial direction; Z_s and its B^T Z +
// Z^T B into d_vxc (FXC_s), Z_z into d_vxcz (FXC_z), with the drop
// screening (z_and_vxc); symmetrised.
int enqueue_fxc_uks(xcq_context_s* ctx, int func) {
  const int gga = (func == XCQ_FUNC_PBE || func == XCQ_FUNC_PBE0) ? 1 : 0;
  const int mode = derivative_mode(ctx);
  CtxView v = view_of(ctx);
  v.in_stage = mode;
  v.tmax = nullptr;
  v.vxc_drop = 0.0;
  hipStream_t s = ctx->stream;
  const size_t blk = (size_t)256 * ctx->pblocks;
  double* dd[4] = {ctx->d_fxcu, ctx->d_fxcu + blk, ctx->d_fxcu + 2 * blk, ctx->d_fxcu + 3 * blk};
  double* zs = ctx->d_fxcu + 4 * blk;
  double* zz = ctx->d_fxcu + 5 * blk;
  const double* mats[4] = {ctx->d_P, ctx->d_Pz, ctx->d_tP, ctx->d_tPz};
  int rc;
  if ((rc = launch_fill(s, ctx->d_vxc, ctx->p_total)) ||
      (rc = launch_fill(s, ctx->d_vxcz, ctx->p_total)))
    return rc;
  if (ctx->ncounters &&
      (rc = launch_fill(s, reinterpret_cast<double*>(ctx->d_counters), ctx->ncounters)))
    return rc;
  for (const Chunk& c : ctx->chunks) {
    const WorkItem* cw = v.work + c.coll.off;
    if ((rc = launch_collocation_fp64(s, cw, c.coll.n, v.tasks, v.shells, v.shell_list, v.lsh,
                                      v.px, v.py, v.pz, static_cast<double*>(v.scratch),
                                      gga && mode == 0, ctx->max_l)))
      return rc;
    for (int pass = 0; pass < 4; ++pass)
      for (int k = 0; k < 4; ++k)
        if ((rc = launch_density(v, c.dens[k], k == 0 ? 32 : 64, false, XCQ_PREC_FP64, func, gga,
                                 1.0, 0, mats[pass], dd[pass])))
          return rc;
    if ((rc = launch_fxc_points_uks(v, c.coll, func, gga, dd[0], dd[1], dd[2], dd[3], zs, zz)))
      return rc;
    for (int comp = 0; comp < 2; ++comp) {
      CtxView vc = v;
      vc.vxc = comp == 0 ? ctx->d_vxc : ctx->d_vxcz;
      if ((rc = z_and_vxc(ctx, vc, c, mode, gga, comp == 0 ? zs : zz))) return rc;
    }
  }
  for (double* m : {ctx->d_vxc, ctx->d_vxcz})
    if ((rc = launch_symmetrize(s, 1, ctx->max_nbf, ctx->d_mol_n, ctx->d_mol_off, m))) return rc;
  return launch_reduce_sums(s, 1, ctx->d_mol_blk, ctx->d_red, ctx->d_rpart, nullptr, 1.0,
                            ctx->d_out2);
}

}  // namespace

int xcq_upload_trial_density_z(xcq_context ctx, int32_t nbf, const double* tPz, int64_t ldp) {
  XCQ_REQUIRE(ctx && tPz, XCQ_ERROR_INVALID_ARGUMENT, "null argument");
  XCQ_REQUIRE(ctx->mols.size() == 1, XCQ_ERROR_STATE,
              "the FXC contraction needs a single-molecule context");
  const MolInfo& M = ctx->mols[0];
  XCQ_REQUIRE(nbf == M.nbf && nbf > 0, XCQ_ERROR_INVALID_ARGUMENT,
              "density dimension does not match the basis");
  XCQ_REQUIRE(ldp >= nbf, XCQ_ERROR_INVALID_ARGUMENT, "ldp < nbf");
  XCQ_HIP(hipSetDevice(ctx->device));
  int rc;
  if (!ctx->d_tPz && (rc = dalloc(&ctx->d_tPz, (size_t)ctx->p_total))) return rc;
  XCQ_HIP(hipMemcpy2DAsync(ctx->d_tPz, (size_t)nbf * 8, tPz, (size_t)ldp * 8, (size_t)nbf * 8,
                           nbf, hipMemcpyHostToDevice, ctx->stream));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  ctx->have_tpz = true;
  return XCQ_SUCCESS;
}

int xcq_fxc_contraction_uks(xcq_context ctx, int functional) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(functional == XCQ_FUNC_SVWN5 || functional == XCQ_FUNC_PBE ||
                  functional == XCQ_FUNC_PBE0,
              XCQ_ERROR_INVALID_ARGUMENT, "unknown functional");
  XCQ_REQUIRE(ctx->ntasks > 0, XCQ_ERROR_STATE, "no tasks uploaded");
  XCQ_REQUIRE(ctx->mols.size() == 1, XCQ_ERROR_STATE,
              "the FXC contraction needs a single-molecule context");
  XCQ_REQUIRE(ctx->mols[0].have_density && ctx->have_pz && ctx->have_tp && ctx->have_tpz,
              XCQ_ERROR_STATE,
              "the UKS FXC contraction needs P_s, P_z (xcq_upload_density, "
              "xcq_upload_density_z) and tP_s, tP_z (xcq_upload_trial_density, "
              "xcq_upload_trial_density_z)");
  XCQ_HIP(hipSetDevice(ctx->device));
  int rc;
  if (!ctx->d_fxcu &&
      (rc = dalloc(&ctx->d_fxcu, (size_t)6 * 256 * std::max<int64_t>(ctx->pblocks, 1))))
    return rc;
  if (!ctx->d_vxcz && (rc = dalloc(&ctx->d_vxcz, (size_t)ctx->p_total))) return rc;
  ctx->last_want_vxc = 0;
  ctx->last_vxc_eval = 0;
  ctx->last_graph = false;
  ctx->last_uks = false;
  ctx->last_fxc = false;
  ctx->last_grad = false;
  ctx->last_fxc_uks = true;
  return enqueue_fxc_uks(ctx, functional);
}

int xcq_download_fxc_uks(xcq_context ctx, double* nel, double* fxc_s, double* fxc_z,
                         int64_t ldfxc) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(ctx->last_fxc_uks, XCQ_ERROR_STATE,
              "the last evaluation was not a UKS FXC contraction");
  const MolInfo& M = ctx->mols[0];
  XCQ_HIP(hipSetDevice(ctx->device));
  double out[2];
  XCQ_HIP(hipMemcpyAsync(out, ctx->d_out2, 16, hipMemcpyDeviceToHost, ctx->stream));
  for (int c = 0; c < 2; ++c) {
    double* dst = c == 0 ? fxc_s : fxc_z;
    if (!dst) continue;
    XCQ_REQUIRE(ldfxc >= M.nbf, XCQ_ERROR_INVALID_ARGUMENT, "ldfxc < nbf");
    XCQ_HIP(hipMemcpy2DAsync(dst, (size_t)ldfxc * 8, c == 0 ? ctx->d_vxc : ctx->d_vxcz,
                             (size_t)M.nbf * 8, (size_t)M.nbf * 8, M.nbf,
                             hipMemcpyDeviceToHost, ctx->stream));
  }
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  if (nel) *nel = out[1];
  return XCQ_SUCCESS;
}

int xcq_download_fxc(xcq_context ctx, double* nel, double* fxc, int64_t ldfxc) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(ctx->last_fxc, XCQ_ERROR_STATE, "the last evaluation was not an FXC contraction");
  const MolInfo& M = ctx->mols[0];
  XCQ_HIP(hipSetDevice(ctx->device));
  double out[2];
  XCQ_HIP(hipMemcpyAsync(out, ctx->d_out2, 16, hipMemcpyDeviceToHost, ctx->stream));
  if (fxc) {
    XCQ_REQUIRE(ldfxc >= M.nbf, XCQ_ERROR_INVALID_ARGUMENT, "ldfxc < nbf");
    XCQ_HIP(hipMemcpy2DAsync(fxc, (size_t)ldfxc * 8, ctx->d_vxc, (size_t)M.nbf * 8,
                             (size_t)M.nbf * 8, M.nbf, hipMemcpyDeviceToHost, ctx->stream));
  }
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  if (nel) *nel = out[1];
  return XCQ_SUCCESS;
}

// ---- EXC gradient (GauXC eval_exc_grad(P)) ----------------------------------
namespace {

// One RKS EXC gradient on the context's stream (FP64): per chunk,
// collocation; the density stage (factor 2) leaving raw per-point sums and,
// for GGAs, its X = 2 P_sub B in a plane; per-point Za coefficients and
// w eps rho; Za; Y = 2 P_sub Za into plane 2; the basis-function part and
// the grid-weight part into d_grad (the latter only with weight_derivs).
// Up to f functions the basis is evaluated in-stage (collocation stores the
// value plane only; X in plane 1, Za by k_z_in_stage into plane 4); with g
// functions on stored planes (B and grad B; X in plane 4, Za over grad_z B
// in plane 3).
int enqueue_grad(xcq_context_s* ctx, int func, int weight_derivs) {
  const int gga = (func == XCQ_FUNC_PBE || func == XCQ_FUNC_PBE0) ? 1 : 0;
  const int mode = derivative_mode(ctx);
  CtxView v = view_of(ctx);
  v.in_stage = mode;
  v.tmax = nullptr;
  v.vxc_drop = 0.0;
  hipStream_t s = ctx->stream;
  const size_t blk = (size_t)256 * ctx->pblocks;
  double* dp = ctx->d_uks;
  double* coef = dp + blk;
  double* wf = coef + blk;
  int rc;
  if ((rc = launch_fill(s, ctx->d_grad, 3 * (int64_t)ctx->grad_natoms))) return rc;
  if (ctx->ncounters &&
      (rc = launch_fill(s, reinterpret_cast<double*>(ctx->d_counters), ctx->ncounters)))
    return rc;
  for (const Chunk& c : ctx->chunks) {
    const WorkItem* cw = v.work + c.coll.off;
    if ((rc = launch_collocation_fp64(s, cw, c.coll.n, v.tasks, v.shells, v.shell_list, v.lsh,
                                      v.px, v.py, v.pz, static_cast<double*>(v.scratch),
                                      gga && mode == 0, ctx->max_l)))
      return rc;
    // X: plane 1 in the in-stage mode (no gradient planes are stored),
    // plane 4 on stored planes.
    const int x_pl = mode == 3 ? 1 : 4;
    for (int k = 0; k < 4; ++k)
      if ((rc = launch_density(v, c.dens[k], k == 0 ? 32 : 64, false, XCQ_PREC_FP64, func, gga,
                               2.0, 0, ctx->d_P, dp, gga ? x_pl : -1)))
        return rc;
    if ((rc = launch_grad_points(v, c.coll, func, gga, dp, coef, wf))) return rc;
    // Za: plane 4 (in-stage evaluation) or over grad_z B in plane 3.
    const int za_pl = mode == 3 ? 4 : 3;
    if (mode == 3) rc = z_from_point_coeffs(ctx, v, c, mode, gga, coef);
    else rc = launch_grad_zpair(v, c.coll, gga, coef, nullptr, 3, -1);
    if (rc || (rc = launch_plane_product(v, c.coll, za_pl, 2))) return rc;
    if ((rc = launch_grad_contract(v, c.coll, gga, ctx->d_shell_atom, coef, weight_derivs,
                                   ctx->d_grad, x_pl, 2, ctx->max_l)))
      return rc;
    if (weight_derivs &&
        (rc = launch_ssf_grad(v, c.coll, ctx->grad_natoms, ctx->d_gxyz, ctx->d_dnear, wf,
                              ctx->d_grad, ctx->grad_lists ? ctx->d_gnbr_idx : nullptr,
                              ctx->grad_lists ? ctx->d_gnbr_r : nullptr)))
      return rc;
  }
  return launch_reduce_sums(s, 1, ctx->d_mol_blk, ctx->d_red, ctx->d_rpart, nullptr, 1.0,
                            ctx->d_out2);
}

// One UKS EXC gradient: per chunk, collocation; the density stage twice (P_s,
// then P_z; factor 1); the polarised functional's coefficients written over
// the sums, w eps rho; Y = P_s Za_s + P_z Za_z into plane 1, X_N = P_s B
// into plane 2 and X_Z = P_z B into plane 3 (GGA); the basis-function and
// grid-weight parts.  In-stage mode (up to f functions): X_N and X_Z come
// from the density passes, Za_s and Za_z in turn from k_z_in_stage through
// plane 4.  Stored planes (g functions): Za_s into plane 4 and Za_z over
// grad_z B in plane 3, X_N and X_Z by plane products after Y.
int enqueue_grad_uks(xcq_context_s* ctx, int func, int weight_derivs) {
  const int gga = (func == XCQ_FUNC_PBE || func == XCQ_FUNC_PBE0) ? 1 : 0;
  const int mode = derivative_mode(ctx);
  CtxView v = view_of(ctx);
  v.in_stage = mode;
  v.tmax = nullptr;
  v.vxc_drop = 0.0;
  hipStream_t s = ctx->stream;
  const size_t blk = (size_t)256 * ctx->pblocks;
  double* ds = ctx->d_uks;
  double* dz = ds + blk;
  double* wf = dz + blk;
  int rc;
  if ((rc = launch_fill(s, ctx->d_grad, 3 * (int64_t)ctx->grad_natoms))) return rc;
  if (ctx->ncounters &&
      (rc = launch_fill(s, reinterpret_cast<double*>(ctx->d_counters), ctx->ncounters)))
    return rc;
  for (const Chunk& c : ctx->chunks) {
    const WorkItem* cw = v.work + c.coll.off;
    if ((rc = launch_collocation_fp64(s, cw, c.coll.n, v.tasks, v.shells, v.shell_list, v.lsh,
                                      v.px, v.py, v.pz, static_cast<double*>(v.scratch),
                                      gga && mode == 0, ctx->max_l)))
      return rc;
    // In-stage mode: the density passes leave X_N in plane 2 and X_Z in
    // plane 3 (GGA), which no gradient plane occupies.
    for (int pass = 0; pass < 2; ++pass)
      for (int k = 0; k < 4; ++k)
        if ((rc = launch_density(v, c.dens[k], k == 0 ? 32 : 64, false, XCQ_PREC_FP64, func, gga,
                                 1.0, 0, pass == 0 ? ctx->d_P : ctx->d_Pz,
                                 pass == 0 ? ds : dz,
                                 (gga && mode == 3) ? 2 + pass : -1)))
          return rc;
    if ((rc = launch_grad_points_uks(v, c.coll, func, gga, ds, dz, wf))) return rc;
    if (mode == 3) {
      // Za_s, then Za_z, through plane 4: Y = P_s Za_s + P_z Za_z in plane 1.
      if ((rc = z_from_point_coeffs(ctx, v, c, mode, gga, ds)) ||
          (rc = launch_plane_product(v, c.coll, 4, 1, ctx->d_P, 1.0, 0)) ||
          (rc = z_from_point_coeffs(ctx, v, c, mode, gga, dz)) ||
          (rc = launch_plane_product(v, c.coll, 4, 1, ctx->d_Pz, 1.0, 1)))
        return rc;
    } else {
      if ((rc = launch_grad_zpair(v, c.coll, gga, ds, dz, 4, 3))) return rc;
      if ((rc = launch_plane_product(v, c.coll, 4, 1, ctx->d_P, 1.0, 0)) ||
          (rc = launch_plane_product(v, c.coll, 3, 1, ctx->d_Pz, 1.0, 1)))
        return rc;
      if (gga && ((rc = launch_plane_product(v, c.coll, 0, 2, ctx->d_P, 1.0, 0)) ||
                  (rc = launch_plane_product(v, c.coll, 0, 3, ctx->d_Pz, 1.0, 0))))
        return rc;
    }
    if ((rc = launch_grad_contract_uks(v, c.coll, gga, ctx->d_shell_atom, ds, dz, weight_derivs,
                                       ctx->d_grad, ctx->max_l)))
      return rc;
    if (weight_derivs &&
        (rc = launch_ssf_grad(v, c.coll, ctx->grad_natoms, ctx->d_gxyz, ctx->d_dnear, wf,
                              ctx->d_grad, ctx->grad_lists ? ctx->d_gnbr_idx : nullptr,
                              ctx->grad_lists ? ctx->d_gnbr_r : nullptr)))
      return rc;
  }
  return launch_reduce_sums(s, 1, ctx->d_mol_blk, ctx->d_red, ctx->d_rpart, nullptr, 1.0,
                            ctx->d_out2);
}

// Checks and geometry upload shared by the RKS and UKS gradients.
int grad_prepare(xcq_context_s* ctx, int functional, int32_t natoms, const double* xyz) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(functional == XCQ_FUNC_SVWN5 || functional == XCQ_FUNC_PBE ||
                  functional == XCQ_FUNC_PBE0,
              XCQ_ERROR_INVALID_ARGUMENT, "unknown functional");
  XCQ_REQUIRE(natoms > 0 && xyz, XCQ_ERROR_INVALID_ARGUMENT, "no atoms");
  XCQ_REQUIRE(ctx->ntasks > 0, XCQ_ERROR_STATE, "no tasks uploaded");
  XCQ_REQUIRE(ctx->mols.size() == 1, XCQ_ERROR_STATE,
              "the EXC gradient needs a single-molecule context");
  XCQ_REQUIRE(ctx->mols[0].have_density, XCQ_ERROR_STATE,
              "the EXC gradient needs P (xcq_upload_density)");
  for (const DevTask& d : ctx->h_tasks)
    XCQ_REQUIRE(d.iparent >= 0 && d.iparent < natoms, XCQ_ERROR_INVALID_ARGUMENT,
                "a task's parent atom is outside the atom list");
  XCQ_HIP(hipSetDevice(ctx->device));
  // Shell -> atom: the atom at exactly the shell's centre, -1 if none
  // (BasisSetMap, basisset_map.hpp:69-77).
  const size_t nsh = ctx->sh_size.size();
  std::vector<int32_t> sa(nsh, -1);
  for (size_t i = 0; i < nsh; ++i)
    for (int32_t a = 0; a < natoms && sa[i] < 0; ++a)
      if (xyz[3 * a] == ctx->sh_origin[3 * i] && xyz[3 * a + 1] == ctx->sh_origin[3 * i + 1] &&
          xyz[3 * a + 2] == ctx->sh_origin[3 * i + 2])
        sa[i] = a;
  // R_AB and the nearest-neighbour distance of each atom
  // (MolMeta::dist_nearest; a single atom has none, its points all keep
  // their weights), and the neighbour lists of the weight part's bounded
  // loops (up to kSsfNbrMaxAtoms atoms).
  std::vector<double> rab, dn, hnr;
  std::vector<int32_t> hni;
  ssf_molmeta(natoms, xyz, rab, dn);
  const bool lists = ctx->ssf_lists && natoms <= kSsfNbrMaxAtoms;
  if (lists) ssf_neighbours(natoms, rab, hni, hnr);
  int rc;
  if (natoms != ctx->grad_natoms) {
    dfree(ctx->d_gxyz);
    dfree(ctx->d_dnear);
    dfree(ctx->d_grad);
    dfree(ctx->d_gnbr_idx);
    dfree(ctx->d_gnbr_r);
    ctx->grad_natoms = 0;
    if ((rc = dalloc(&ctx->d_gxyz, 3 * (size_t)natoms)) ||
        (rc = dalloc(&ctx->d_dnear, (size_t)natoms)) ||
        (rc = dalloc(&ctx->d_grad, 3 * (size_t)natoms)))
      return rc;
    ctx->grad_natoms = natoms;
  }
  if (lists && !ctx->d_gnbr_idx &&
      ((rc = dalloc(&ctx->d_gnbr_idx, hni.size())) || (rc = dalloc(&ctx->d_gnbr_r, hnr.size()))))
    return rc;
  ctx->grad_lists = lists;
  dfree(ctx->d_shell_atom);
  if ((rc = dalloc(&ctx->d_shell_atom, std::max<size_t>(nsh, 1)))) return rc;
  XCQ_HIP(hipMemcpy(ctx->d_gxyz, xyz, 24 * (size_t)natoms, hipMemcpyHostToDevice));
  XCQ_HIP(hipMemcpy(ctx->d_dnear, dn.data(), 8 * (size_t)natoms, hipMemcpyHostToDevice));
  XCQ_HIP(hipMemcpy(ctx->d_shell_atom, sa.data(), 4 * nsh, hipMemcpyHostToDevice));
  if (lists) {
    XCQ_HIP(hipMemcpy(ctx->d_gnbr_idx, hni.data(), 4 * hni.size(), hipMemcpyHostToDevice));
    XCQ_HIP(hipMemcpy(ctx->d_gnbr_r, hnr.data(), 8 * hnr.size(), hipMemcpyHostToDevice));
  }
  if ((rc = ensure_point_buffers(ctx))) return rc;
  ctx->last_want_vxc = 0;
  ctx->last_vxc_eval = 0;
  ctx->last_graph = false;
  ctx->last_uks = false;
  ctx->last_fxc = false;
  ctx->last_fxc_uks = false;
  ctx->last_grad = true;
  return XCQ_SUCCESS;
}

}  // namespace

int xcq_exc_grad(xcq_context ctx, int functional, int32_t natoms, const double* xyz,
                 int weight_derivs) {
  const int rc = grad_prepare(ctx, functional, natoms, xyz);
  return rc ? rc : enqueue_grad(ctx, functional, weight_derivs ? 1 : 0);
}

int xcq_exc_grad_uks(xcq_context ctx, int functional, int32_t natoms, const double* xyz,
                     int weight_derivs) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(ctx->have_pz, XCQ_ERROR_STATE,
              "the UKS EXC gradient needs P_z (xcq_upload_density_z)");
  const int rc = grad_prepare(ctx, functional, natoms, xyz);
  return rc ? rc : enqueue_grad_uks(ctx, functional, weight_derivs ? 1 : 0);
}

int xcq_download_exc_grad(xcq_context ctx, double* exc, double* nel, double* grad) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(ctx->last_grad, XCQ_ERROR_STATE, "the last evaluation was not an EXC gradient");
  XCQ_HIP(hipSetDevice(ctx->device));
  double out[2];
  XCQ_HIP(hipMemcpyAsync(out, ctx->d_out2, 16, hipMemcpyDeviceToHost, ctx->stream));
  if (grad)
    XCQ_HIP(hipMemcpyAsync(grad, ctx->d_grad, 24 * (size_t)ctx->grad_natoms,
                           hipMemcpyDeviceToHost, ctx->stream));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  if (exc) *exc = out[0];
  if (nel) *nel = out[1];
  return XCQ_SUCCESS;
}

int xcq_integrate_den(xcq_context ctx, double* nel) {
  XCQ_REQUIRE(nel, XCQ_ERROR_INVALID_ARGUMENT, "null output");
  XCQ_REQUIRE(ctx && ctx->mols.size() == 1, XCQ_ERROR_STATE,
              "integrate_den needs a single-molecule context");
  int rc = run_pipeline(ctx, 0, XCQ_PREC_FP64, 0);
  if (rc) return rc;
  double out[2];
  XCQ_HIP(hipMemcpyAsync(out, ctx->d_out2, 16, hipMemcpyDeviceToHost, ctx->stream));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  *nel = out[1];
  return XCQ_SUCCESS;
}

int xcq_download_results_mol(xcq_context ctx, int32_t mol, double* exc, double* nel,
                             double* vxc, int64_t ldvxc) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(mol >= 0 && mol < (int32_t)ctx->mols.size(), XCQ_ERROR_INVALID_ARGUMENT,
              "molecule index out of range");
  const MolInfo& M = ctx->mols[mol];
  XCQ_HIP(hipSetDevice(ctx->device));
  double out[2];
  XCQ_HIP(hipMemcpyAsync(out, ctx->d_out2 + 2 * mol, 16, hipMemcpyDeviceToHost, ctx->stream));
  if (vxc) {
    XCQ_REQUIRE(ctx->last_want_vxc, XCQ_ERROR_STATE, "last evaluation did not compute VXC");
    XCQ_REQUIRE(ldvxc >= M.nbf, XCQ_ERROR_INVALID_ARGUMENT, "ldvxc < nbf");
    XCQ_HIP(hipMemcpy2DAsync(vxc, (size_t)ldvxc * 8, ctx->d_vxc + M.p_off, (size_t)M.nbf * 8,
                             (size_t)M.nbf * 8, M.nbf, hipMemcpyDeviceToHost, ctx->stream));
  }
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  if (exc) *exc = out[0];
  if (nel) *nel = out[1];
  return XCQ_SUCCESS;
}

int xcq_download_results(xcq_context ctx, double* exc, double* nel, double* vxc, int64_t ldvxc) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(ctx->mols.size() == 1, XCQ_ERROR_STATE,
              "context holds several molecules: use xcq_download_results_mol");
  return xcq_download_results_mol(ctx, 0, exc, nel, vxc, ldvxc);
}

int xcq_stage_times(xcq_context ctx, double* ms, int n) {
  XCQ_REQUIRE(ctx && ms, XCQ_ERROR_INVALID_ARGUMENT, "null argument");
  XCQ_HIP(hipSetDevice(ctx->device));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  double st[kStages + 1] = {0};
  if (ctx->last_graph) {
    // A replayed graph has no per-stage events: report the total only.
    float f = 0.f;
    if (hipEventElapsedTime(&f, ctx->ev_total[0], ctx->ev_total[1]) == hipSuccess)
      st[kStages] = f;
    else
      (void)hipGetLastError();
  } else {
    const int nst = ctx->last_vxc_eval ? 5 : 3;
    for (size_t ci = 0; ci < ctx->chunks.size(); ++ci)
      for (int k = 0; k < nst; ++k) {
        float f = 0.f;
        hipEvent_t* ev = ctx->ev.data() + ci * kStages * 2;
        if (hipEventElapsedTime(&f, ev[2 * k], ev[2 * k + 1]) == hipSuccess) st[k] += f;
        else (void)hipGetLastError();
      }
    if (ctx->last_vxc_eval && !ctx->ev.empty()) {
      float f = 0.f;
      if (hipEventElapsedTime(&f, ctx->ev[10], ctx->ev[11]) == hipSuccess) st[5] = f;
      else (void)hipGetLastError();
    }
    for (int k = 0; k < kStages; ++k) st[kStages] += st[k];
  }
  const int m = std::min(n, kStages + 1);
  for (int k = 0; k < m; ++k) ms[k] = st[k];
  return m;
}

int xcq_work_counters(xcq_context ctx, int64_t* out, int n) {
  XCQ_REQUIRE(ctx && out, XCQ_ERROR_INVALID_ARGUMENT, "null argument");
  const int m = std::min(n, 4);
  for (int k = 0; k < m; ++k) out[k] = ctx->counters[k];
  return m;
}

namespace {

// Device copies of the SSF geometry: atom coordinates, R_AB, nearest
// distances and (up to kSsfNbrMaxAtoms atoms) the neighbour lists.
struct SsfGeometry {
  double *xyz = nullptr, *rab = nullptr, *dn = nullptr, *nr = nullptr;
  int32_t* ni = nullptr;
  ~SsfGeometry() {
    dfree(xyz);
    dfree(rab);
    dfree(dn);
    dfree(nr);
    dfree(ni);
  }
  int upload(int32_t natoms, const double* atom_xyz) {
    std::vector<double> hr, hd, hnr;
    std::vector<int32_t> hni;
    ssf_molmeta(natoms, atom_xyz, hr, hd);
    const bool lists = natoms <= kSsfNbrMaxAtoms;
    if (lists) ssf_neighbours(natoms, hr, hni, hnr);
    int rc;
    if ((rc = dalloc(&xyz, 3 * (size_t)natoms)) || (rc = dalloc(&rab, hr.size())) ||
        (rc = dalloc(&dn, (size_t)natoms)))
      return rc;
    if (lists && ((rc = dalloc(&ni, hni.size())) || (rc = dalloc(&nr, hnr.size())))) return rc;
    XCQ_HIP(hipMemcpy(xyz, atom_xyz, 24 * (size_t)natoms, hipMemcpyHostToDevice));
    XCQ_HIP(hipMemcpy(rab, hr.data(), 8 * hr.size(), hipMemcpyHostToDevice));
    XCQ_HIP(hipMemcpy(dn, hd.data(), 8 * hd.size(), hipMemcpyHostToDevice));
    if (lists) {
      XCQ_HIP(hipMemcpy(ni, hni.data(), 4 * hni.size(), hipMemcpyHostToDevice));
      XCQ_HIP(hipMemcpy(nr, hnr.data(), 8 * hnr.size(), hipMemcpyHostToDevice));
    }
    return XCQ_SUCCESS;
  }
};

}  // namespace

int xcq_ssf_weights(int device, int32_t natoms, const double* atom_xyz, int64_t npts,
                    const double* points, const int32_t* parent, double* weights) {
  XCQ_REQUIRE(natoms > 0 && atom_xyz, XCQ_ERROR_INVALID_ARGUMENT, "empty molecule");
  XCQ_REQUIRE(npts >= 0 && (npts == 0 || (points && parent && weights)),
              XCQ_ERROR_INVALID_ARGUMENT, "bad point arrays");
  for (int64_t i = 0; i < npts; ++i)
    XCQ_REQUIRE(parent[i] >= 0 && parent[i] < natoms, XCQ_ERROR_INVALID_ARGUMENT,
                "parent atom out of range");
  if (npts == 0) return XCQ_SUCCESS;
  int n = 0;
  if (xcq_device_count(&n) != XCQ_SUCCESS || device < 0 || device >= n) {
    set_error("no HIP device available");
    return XCQ_ERROR_NO_DEVICE;
  }
  XCQ_HIP(hipSetDevice(device));
  SsfGeometry g;
  int rc = g.upload(natoms, atom_xyz);
  if (rc) return rc;
  // Points stay interleaved (x, y, z): the kernel reads them with stride 3.
  double *d_pts = nullptr, *d_w = nullptr;
  int32_t* d_par = nullptr;
  auto cleanup = [&] {
    dfree(d_pts);
    dfree(d_w);
    dfree(d_par);
  };
  if ((rc = dalloc(&d_pts, 3 * (size_t)npts)) || (rc = dalloc(&d_w, (size_t)npts)) ||
      (rc = dalloc(&d_par, (size_t)npts))) {
    cleanup();
    return rc;
  }
  hipError_t e = hipMemcpy(d_pts, points, 24 * (size_t)npts, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_w, weights, 8 * (size_t)npts, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_par, parent, 4 * (size_t)npts, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error(std::string("SSF upload: ") + hipGetErrorString(e));
    cleanup();
    return XCQ_ERROR_HIP;
  }
  rc = launch_ssf(nullptr, natoms, g.xyz, g.rab, g.dn, npts, d_pts, d_pts + 1, d_pts + 2, 3,
                  d_par, d_w, nullptr, g.ni, g.nr);
  if (rc == XCQ_SUCCESS) {
    e = hipMemcpy(weights, d_w, 8 * (size_t)npts, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      set_error(std::string("SSF download: ") + hipGetErrorString(e));
      rc = XCQ_ERROR_HIP;
    }
  }
  cleanup();
  return rc;
}

// ---- SSF weights without device allocation ----
// Scratch of xcq_ssf_weights_noalloc: atom coordinates, nearest-neighbour
// distances, a bad-parent flag and the inter-atomic distance matrix of at
// most XCQ_SSF_NOALLOC_MAX_ATOMS atoms, 8.4 MB per device, held by every
// device the library runs on until the process ends.  In host code the name
// `scratch` is only a stand-in: its device address is looked up once per
// device (hipGetSymbolAddress, with that device current) and only that
// address is passed to kernels and copies.
namespace {
constexpr size_t kNoAllocMaxAtoms = XCQ_SSF_NOALLOC_MAX_ATOMS;
constexpr size_t kNoAllocXyz = 0;                                   // 3 * max doubles
constexpr size_t kNoAllocDnear = 3 * kNoAllocMaxAtoms;              // max doubles
constexpr size_t kNoAllocFlag = 4 * kNoAllocMaxAtoms;               // one int32 (32 doubles reserved)
constexpr size_t kNoAllocRab = 4 * kNoAllocMaxAtoms + 32;           // max * max doubles
constexpr size_t kNoAllocDoubles = kNoAllocRab + kNoAllocMaxAtoms * kNoAllocMaxAtoms;
}  // namespace
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on `device` (current), looked up on first use.
int noalloc_scratch(int device, double** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return XCQ_ERROR_HIP;
    }
    g_scratch_addr[device] = static_cast<double*>(p);
  }
  *out = g_scratch_addr[device];
  return XCQ_SUCCESS;
}
}  // namespace

int xcq_ssf_weights_noalloc(int device, int32_t natoms, const double* atom_xyz, int64_t npts,
                            const double* points, const int32_t* parent, double* weights) {
  XCQ_REQUIRE(natoms > 0 && atom_xyz, XCQ_ERROR_INVALID_ARGUMENT, "empty molecule");
  XCQ_REQUIRE((size_t)natoms <= kNoAllocMaxAtoms, XCQ_ERROR_INVALID_ARGUMENT,
              "xcq_ssf_weights_noalloc: natoms exceeds XCQ_SSF_NOALLOC_MAX_ATOMS (" +
                  std::to_string(kNoAllocMaxAtoms) + ")");
  XCQ_REQUIRE(npts >= 0 && (npts == 0 || (points && parent && weights)),
              XCQ_ERROR_INVALID_ARGUMENT, "bad point arrays");
  if (npts == 0) return XCQ_SUCCESS;
  int n = 0;
  if (xcq_device_count(&n) != XCQ_SUCCESS || device < 0 || device >= n ||
      device >= kMaxDevices) {
    set_error("no HIP device available");
    return XCQ_ERROR_NO_DEVICE;
  }
  XCQ_HIP(hipSetDevice(device));
  double* base = nullptr;
  int rc = noalloc_scratch(device, &base);
  if (rc) return rc;
  double* d_xyz = base + kNoAllocXyz;
  double* d_dn = base + kNoAllocDnear;
  int32_t* d_flag = reinterpret_cast<int32_t*>(base + kNoAllocFlag);
  double* d_rab = base + kNoAllocRab;
  XCQ_HIP(hipMemcpy(d_xyz, atom_xyz, 24 * (size_t)natoms, hipMemcpyHostToDevice));
  if ((rc = launch_molmeta(nullptr, natoms, d_xyz, d_rab, d_dn, d_flag))) return rc;
  if ((rc = launch_ssf(nullptr, natoms, d_xyz, d_rab, d_dn, npts, points, points + 1, points + 2,
                       3, parent, weights, d_flag)))
    return rc;
  int32_t bad = 0;
  XCQ_HIP(hipMemcpy(&bad, d_flag, sizeof(bad), hipMemcpyDeviceToHost));
  XCQ_REQUIRE(bad == 0, XCQ_ERROR_INVALID_ARGUMENT,
              "parent atom out of range (the weights of those points are unchanged)");
  return XCQ_SUCCESS;
}

int xcq_ssf_weights_tasks(xcq_context ctx, int32_t natoms, const double* atom_xyz) {
  XCQ_REQUIRE(ctx && ctx->ntasks > 0, XCQ_ERROR_STATE, "no tasks uploaded");
  XCQ_REQUIRE(ctx->mols.size() == 1, XCQ_ERROR_STATE,
              "SSF on resident tasks needs a single-molecule context");
  XCQ_REQUIRE(natoms > 0 && atom_xyz, XCQ_ERROR_INVALID_ARGUMENT, "empty molecule");
  XCQ_HIP(hipSetDevice(ctx->device));
  std::vector<int32_t> parent(ctx->npts_total);
  for (const auto& d : ctx->h_tasks) {
    XCQ_REQUIRE(d.iparent >= 0 && d.iparent < natoms, XCQ_ERROR_INVALID_ARGUMENT,
                "task parent atom out of range");
    for (int32_t i = 0; i < d.npts; ++i) parent[d.pt_off + i] = d.iparent;
  }
  SsfGeometry g;
  int rc = g.upload(natoms, atom_xyz);
  if (rc) return rc;
  int32_t* d_par = nullptr;
  if ((rc = dalloc(&d_par, parent.size()))) return rc;
  hipError_t e = hipMemcpy(d_par, parent.data(), 4 * parent.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    rc = launch_ssf(ctx->stream, natoms, g.xyz, g.rab, g.dn, ctx->npts_total, ctx->d_px,
                    ctx->d_py, ctx->d_pz, 1, d_par, ctx->d_w, nullptr, g.ni, g.nr);
    if (rc == XCQ_SUCCESS) e = hipStreamSynchronize(ctx->stream);
  }
  dfree(d_par);
  if (e != hipSuccess) {
    set_error(std::string("SSF on resident tasks: ") + hipGetErrorString(e));
    return XCQ_ERROR_HIP;
  }
  return rc;
}

int xcq_download_weights(xcq_context ctx, double* weights) {
  XCQ_REQUIRE(ctx && weights, XCQ_ERROR_INVALID_ARGUMENT, "null argument");
  XCQ_HIP(hipSetDevice(ctx->device));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->npts_total)
    XCQ_HIP(hipMemcpy(weights, ctx->d_w, 8 * (size_t)ctx->npts_total, hipMemcpyDeviceToHost));
  return XCQ_SUCCESS;
}

int xcq_task_collocation(xcq_context ctx, int32_t task, int deriv, int precision, double* out) {
  XCQ_REQUIRE(ctx && out, XCQ_ERROR_INVALID_ARGUMENT, "null argument");
  XCQ_REQUIRE(task >= 0 && task < ctx->ntasks, XCQ_ERROR_INVALID_ARGUMENT,
              "task index out of range");
  XCQ_REQUIRE(precision == XCQ_PREC_FP64 || precision == XCQ_PREC_MIXED,
              XCQ_ERROR_INVALID_ARGUMENT, "unknown precision");
  XCQ_HIP(hipSetDevice(ctx->device));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  const DevTask& d = ctx->h_tasks[task];
  const int nmb = d.npts_pad / kPtBlock;
  std::vector<WorkItem> items(nmb);
  for (int pb = 0; pb < nmb; ++pb) items[pb] = WorkItem{task, pb, 0, 0};
  WorkItem* d_items = nullptr;
  int rc = dalloc(&d_items, items.size());
  if (rc) return rc;
  const int ncomp = deriv ? 4 : 1;
  const int64_t plane = (int64_t)d.nbe * d.npts_pad;
  std::vector<double> host((size_t)ncomp * plane);
  hipError_t e = nmb ? hipMemcpy(d_items, items.data(), items.size() * sizeof(WorkItem),
                                 hipMemcpyHostToDevice)
                     : hipSuccess;
  if (e == hipSuccess && nmb > 0) {
    if (precision == XCQ_PREC_FP64) {
      double* scr = static_cast<double*>(ctx->d_scratch);
      rc = launch_collocation_fp64(ctx->stream, d_items, nmb, ctx->d_tasks, ctx->d_shells,
                                   ctx->d_shell_list, ctx->d_lsh, ctx->d_px, ctx->d_py,
                                   ctx->d_pz, scr, deriv, ctx->max_l);
      if (rc == XCQ_SUCCESS) e = hipStreamSynchronize(ctx->stream);
      if (rc == XCQ_SUCCESS && e == hipSuccess)
        e = hipMemcpy(host.data(), scr + d.b_off, 8 * host.size(), hipMemcpyDeviceToHost);
    } else {
      // FP32 collocation: same layout in float inside the scratch.
      float* scr = static_cast<float*>(ctx->d_scratch);
      rc = launch_collocation_fp32(ctx->stream, d_items, nmb, ctx->d_tasks, ctx->d_shells,
                                   ctx->d_shell_list, ctx->d_lsh, ctx->d_px, ctx->d_py,
                                   ctx->d_pz, scr, deriv, ctx->max_l);
      if (rc == XCQ_SUCCESS) e = hipStreamSynchronize(ctx->stream);
      std::vector<float> hf(host.size());
      if (rc == XCQ_SUCCESS && e == hipSuccess)
        e = hipMemcpy(hf.data(), scr + d.b_off_mx, 4 * hf.size(), hipMemcpyDeviceToHost);
      for (size_t k = 0; k < hf.size(); ++k) host[k] = hf[k];
    }
  }
  dfree(d_items);
  if (rc) return rc;
  if (e != hipSuccess) {
    set_error(std::string("collocation: ") + hipGetErrorString(e));
    return XCQ_ERROR_HIP;
  }
  // [comp][nbe][npts_pad] -> [comp][nbe][npts]
  for (int cpt = 0; cpt < ncomp; ++cpt)
    for (int j = 0; j < d.nbe; ++j)
      std::memcpy(out + ((size_t)cpt * d.nbe + j) * d.npts,
                  host.data() + (size_t)cpt * plane + (size_t)j * d.npts_pad, 8 * (size_t)d.npts);
  return XCQ_SUCCESS;
}

int xcq_eval_functional(int device, int functional, int64_t n, const double* rho,
                        const double* sigma, double* eps, double* vrho, double* vsigma) {
  XCQ_REQUIRE(functional == XCQ_FUNC_SVWN5 || functional == XCQ_FUNC_PBE ||
                  functional == XCQ_FUNC_PBE0,
              XCQ_ERROR_INVALID_ARGUMENT, "unknown functional");
  XCQ_REQUIRE(n >= 0 && (n == 0 || (rho && eps && vrho)), XCQ_ERROR_INVALID_ARGUMENT,
              "bad arrays");
  const bool gga = functional != XCQ_FUNC_SVWN5;
  XCQ_REQUIRE(!gga || n == 0 || (sigma && vsigma), XCQ_ERROR_INVALID_ARGUMENT,
              "GGA needs sigma and vsigma");
  if (n == 0) return XCQ_SUCCESS;
  int nd = 0;
  if (xcq_device_count(&nd) != XCQ_SUCCESS || device < 0 || device >= nd) {
    set_error("no HIP device available");
    return XCQ_ERROR_NO_DEVICE;
  }
  XCQ_HIP(hipSetDevice(device));
  double* buf = nullptr;
  int rc = dalloc(&buf, (size_t)5 * n);
  if (rc) return rc;
  double *dr = buf, *ds = buf + n, *de = buf + 2 * n, *dv = buf + 3 * n, *dw = buf + 4 * n;
  hipError_t e = hipMemcpy(dr, rho, 8 * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess && gga) e = hipMemcpy(ds, sigma, 8 * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    rc = launch_functional_points(nullptr, functional, n, dr, gga ? ds : nullptr, de, dv,
                                  gga ? dw : nullptr);
    if (rc == XCQ_SUCCESS) {
      e = hipMemcpy(eps, de, 8 * (size_t)n, hipMemcpyDeviceToHost);
      if (e == hipSuccess) e = hipMemcpy(vrho, dv, 8 * (size_t)n, hipMemcpyDeviceToHost);
      if (e == hipSuccess && gga) e = hipMemcpy(vsigma, dw, 8 * (size_t)n, hipMemcpyDeviceToHost);
    }
  }
  dfree(buf);
  if (rc) return rc;
  if (e != hipSuccess) {
    set_error(std::string("functional: ") + hipGetErrorString(e));
    return XCQ_ERROR_HIP;
  }
  return XCQ_SUCCESS;
}

}  // extern "C"
