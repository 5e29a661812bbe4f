// C ABI of the device layer (include/gauxc_amd/device.h): context, resident
// data, chunk planning and the stage pipeline of one EXC/VXC evaluation.
//
// HBM layout (all resident until replaced):
//   shells      DevShell[nshells]
//   px/py/pz/w  SoA points and weights of all tasks, in upload order
//   shell list, local->global basis map, local shell -> first function
//   work lists  (task, point block) and SYR2K tiles per chunk
//   P, VXC      nbf x nbf, column-major, ld = nbf
//   red         (w eps rho, w rho) per point block -> deterministic sums
// Chunk scratch (reused by every chunk): per task [5][nbe][npts_pad]
// (B, dB/dx, dB/dy, dB/dz, Z) and the Z coefficients.  Tasks are
// grouped greedily into chunks that fit the arena; each chunk is one launch
// per stage, so every launch carries thousands of workgroups.
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace xcq {

static thread_local std::string g_last_error;
void set_error(const std::string& m) { g_last_error = m; }

int launch_collocation_fp32(hipStream_t s, const Chunk& c, const WorkItem* pblk,
                            const DevTask* tasks, const DevShell* shells,
                            const int32_t* shell_list, const int32_t* lsh_first,
                            const double* px, const double* py, const double* pz, float* scratch,
                            int deriv);
int launch_mixed_pipeline(struct ::xcq_context_s* ctx, const Chunk& c, int func, int gga,
                          int want_vxc, hipEvent_t* ev);

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
}  // namespace

struct xcq_context_s {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t arena_bytes = 0;  // user limit (0 = derive from free memory)
  double fill = 0.9;

  // basis
  DevShell* d_shells = nullptr;
  int32_t nshells = 0, nbf = 0;
  std::vector<int32_t> sh_size, sh_first;

  // tasks
  int32_t ntasks = 0;
  int64_t npts_total = 0, total_pblocks = 0;
  std::vector<DevTask> h_tasks;
  DevTask* d_tasks = nullptr;
  double *d_px = nullptr, *d_py = nullptr, *d_pz = nullptr, *d_w = nullptr;
  int32_t *d_shell_list = nullptr, *d_bfmap = nullptr, *d_lsh = nullptr;
  WorkItem* d_work = nullptr;
  std::vector<Chunk> chunks;

  // density / results
  double* d_P = nullptr;
  double* d_vxc = nullptr;
  double* d_red = nullptr;
  double* d_out2 = nullptr;
  bool have_density = false;
  int last_want_vxc = 0;

  // chunk scratch
  void* d_scratch = nullptr;
  size_t scratch_bytes = 0;
  int64_t max_nq = 0, max_scratch_elems = 0;

  // packed bfloat16 slices of P for the mixed-precision contractions
  void* d_Phi = nullptr;

  // timing
  std::vector<hipEvent_t> ev;  // [chunk][stage][2]
  double stage_ms[kStages + 1] = {0};
  int64_t counters[4] = {0, 0, 0, 0};

  ~xcq_context_s() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto e : ev) (void)hipEventDestroy(e);
    dfree(d_shells);
    dfree(d_tasks);
    dfree(d_px);
    dfree(d_py);
    dfree(d_pz);
    dfree(d_w);
    dfree(d_shell_list);
    dfree(d_bfmap);
    dfree(d_lsh);
    dfree(d_work);
    dfree(d_P);
    dfree(d_vxc);
    dfree(d_red);
    dfree(d_out2);
    if (d_scratch) (void)hipFree(d_scratch);
    if (d_Phi) (void)hipFree(d_Phi);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace xcq {
CtxView view_of(xcq_context_s* c) {
  return CtxView{c->stream, c->d_tasks, c->d_shells, c->d_shell_list, c->d_bfmap, c->d_lsh,
                 c->d_px,   c->d_py,    c->d_pz,     c->d_w,          c->d_work,  c->d_P,
                 c->d_vxc,  c->nbf,     c->d_red,    c->d_scratch,
                 c->scratch_bytes, &c->d_Phi};
}
}  // namespace xcq

namespace {
// The command processor hands workgroup i of a launch to XCD i % 8.  Work
// lists are built task by task, so neighbouring items share a task's B/Z
// panels and P_sub gathers; this permutation gives each XCD runs of kXcdRun
// consecutive items (one L2 then sees all of a run's reuse) while keeping
// the list's large-first order at the granularity of 8 * kXcdRun items.
// The final partial group keeps its natural order.
constexpr size_t kXcds = 8, kXcdRun = 16;
void xcd_interleave(std::vector<xcq::WorkItem>& v, size_t b, size_t e) {
  constexpr size_t G = kXcds * kXcdRun;
  const size_t n = e - b, full = n / G * G;
  if (full == 0) return;
  std::vector<xcq::WorkItem> tmp(v.begin() + b, v.begin() + b + full);
  for (size_t i = 0; i < full; ++i) {
    const size_t g = i / G, o = i % G, x = o / kXcdRun, q = o % kXcdRun;
    v[b + g * G + q * kXcds + x] = tmp[i];
  }
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
  XCQ_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  int rc = upload_sph_table(ctx->stream);
  if (rc) return rc;
  if ((rc = dalloc(&ctx->d_out2, 2))) return rc;
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  *out = ctx.release();
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

int xcq_upload_basis(xcq_context ctx, int32_t nshells, const xcq_shell* shells) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(nshells > 0 && shells, XCQ_ERROR_INVALID_ARGUMENT, "empty basis");
  XCQ_HIP(hipSetDevice(ctx->device));
  std::vector<DevShell> h(nshells);
  ctx->sh_size.assign(nshells, 0);
  ctx->sh_first.assign(nshells + 1, 0);
  for (int32_t i = 0; i < nshells; ++i) {
    const auto& s = shells[i];
    XCQ_REQUIRE(s.l >= 0 && s.l <= kMaxL, XCQ_ERROR_INVALID_ARGUMENT,
                "shell angular momentum out of range (0..4)");
    XCQ_REQUIRE(s.nprim > 0 && s.nprim <= 32, XCQ_ERROR_INVALID_ARGUMENT,
                "shell primitive count out of range (1..32)");
    auto& d = h[i];
    std::memset(&d, 0, sizeof(d));
    for (int k = 0; k < s.nprim; ++k) {
      d.alpha[k] = s.alpha[k];
      d.coeff[k] = s.coeff[k];
    }
    for (int k = 0; k < 3; ++k) d.O[k] = s.origin[k];
    d.nprim = s.nprim;
    d.l = s.l;
    d.pure = s.pure;
    d.size = s.pure ? 2 * s.l + 1 : (s.l + 1) * (s.l + 2) / 2;
    ctx->sh_size[i] = d.size;
    ctx->sh_first[i + 1] = ctx->sh_first[i] + d.size;
  }
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  dfree(ctx->d_shells);
  int rc = dalloc(&ctx->d_shells, nshells);
  if (rc) return rc;
  XCQ_HIP(hipMemcpy(ctx->d_shells, h.data(), sizeof(DevShell) * nshells, hipMemcpyHostToDevice));
  ctx->nshells = nshells;
  const int32_t nbf = ctx->sh_first[nshells];
  if (nbf != ctx->nbf) {
    dfree(ctx->d_P);
    dfree(ctx->d_vxc);
    if (ctx->d_Phi) {
      (void)hipFree(ctx->d_Phi);
      ctx->d_Phi = nullptr;
    }
    ctx->nbf = nbf;
    if ((rc = dalloc(&ctx->d_P, (size_t)nbf * nbf))) return rc;
    if ((rc = dalloc(&ctx->d_vxc, (size_t)nbf * nbf))) return rc;
    ctx->have_density = false;
  }
  ctx->ntasks = 0;  // tasks refer to shell indices: require a new upload
  return XCQ_SUCCESS;
}

int xcq_upload_tasks(xcq_context ctx, int32_t ntasks, const xcq_task* tasks, int64_t npts_total,
                     const double* points, const double* weights, int64_t nshell_entries,
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

  // Validate tasks and build the per-task device descriptors.
  std::vector<DevTask> ht(ntasks);
  std::vector<int32_t> bfmap, lsh;
  int64_t pblocks = 0;
  for (int32_t i = 0; i < ntasks; ++i) {
    const auto& t = tasks[i];
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
    d.npts_pad = (t.npts + kTile - 1) / kTile * kTile;
    d.pt_off = t.point_offset;
    d.sh_off = t.shell_offset;
    d.bf_off = (int64_t)bfmap.size();
    d.lsh_off = (int64_t)lsh.size();
    int32_t nbe = 0, prev = -1;
    for (int32_t k = 0; k < t.nshells; ++k) {
      const int32_t s = shell_lists[t.shell_offset + k];
      XCQ_REQUIRE(s >= 0 && s < ctx->nshells, XCQ_ERROR_INVALID_ARGUMENT,
                  "shell index out of range");
      XCQ_REQUIRE(s > prev, XCQ_ERROR_INVALID_ARGUMENT, "shell lists must be ascending");
      prev = s;
      lsh.push_back(nbe);
      for (int32_t f = 0; f < ctx->sh_size[s]; ++f) bfmap.push_back(ctx->sh_first[s] + f);
      nbe += ctx->sh_size[s];
    }
    XCQ_REQUIRE(t.nbe == nbe, XCQ_ERROR_INVALID_ARGUMENT,
                "task nbe does not match its shell list");
    d.nbe = nbe;
    d.ntile = (nbe + kTile - 1) / kTile;
    d.red_off = pblocks;
    pblocks += d.npts_pad / kPtBlock;
  }

  // Chunk planning against the arena.
  size_t free_b = 0, total_b = 0;
  XCQ_HIP(hipMemGetInfo(&free_b, &total_b));
  const size_t resident = (size_t)npts_total * 32 + bfmap.size() * 4 + lsh.size() * 4 +
                          (size_t)nshell_entries * 4 + (size_t)ntasks * sizeof(DevTask) +
                          (size_t)pblocks * 16 * 8;
  size_t budget;
  if (ctx->arena_bytes) {
    budget = ctx->arena_bytes;
  } else {
    const double avail = ctx->fill * (double)free_b - (double)resident;
    budget = (size_t)std::max(0.0, std::min(avail * 0.5, 48.0 * (1ull << 30)));
  }
  auto task_bytes = [](const DevTask& d) {
    return (size_t)5 * d.nbe * d.npts_pad * 8 + (size_t)4 * d.npts_pad * 8;
  };
  std::vector<Chunk> chunks;
  std::vector<WorkItem> pbl, g2;
  {
    int32_t i = 0;
    while (i < ntasks) {
      Chunk c;
      std::memset(&c, 0, sizeof(c));
      c.task_begin = i;
      size_t used = 0;
      int64_t se = 0, nq = 0;
      c.pblk_off = (int64_t)pbl.size();
      c.g2_off = (int64_t)g2.size();
      while (i < ntasks) {
        auto& d = ht[i];
        const size_t b = task_bytes(d);
        if (used + b > budget) {
          XCQ_REQUIRE(i > c.task_begin, XCQ_ERROR_OUT_OF_MEMORY,
                      "Insufficient memory to even start!");
          break;
        }
        used += b;
        d.b_off = se;
        d.q_off = nq;
        se += (int64_t)5 * d.nbe * d.npts_pad;
        nq += d.npts_pad;
        const int nmb = d.npts_pad / kPtBlock;
        for (int pb = 0; pb < nmb; ++pb) pbl.push_back(WorkItem{i, pb, 0, 0});
        for (int it = 0; it < d.ntile; ++it)
          for (int jt = 0; jt <= it; ++jt) g2.push_back(WorkItem{i, it, jt, 0});
        ++i;
      }
      c.task_end = i;
      c.scratch_elems = se;
      c.nq = nq;
      c.npblk = (int64_t)pbl.size() - c.pblk_off;
      c.ng2 = (int64_t)g2.size() - c.g2_off;
      xcd_interleave(pbl, (size_t)c.pblk_off, pbl.size());
      xcd_interleave(g2, (size_t)c.g2_off, g2.size());
      chunks.push_back(c);
    }
  }

  // Release the previous task set and allocate the new one.
  dfree(ctx->d_tasks);
  dfree(ctx->d_px);
  dfree(ctx->d_py);
  dfree(ctx->d_pz);
  dfree(ctx->d_w);
  dfree(ctx->d_shell_list);
  dfree(ctx->d_bfmap);
  dfree(ctx->d_lsh);
  dfree(ctx->d_work);
  dfree(ctx->d_red);
  if (ctx->d_scratch) {
    (void)hipFree(ctx->d_scratch);
    ctx->d_scratch = nullptr;
  }
  for (auto e : ctx->ev) (void)hipEventDestroy(e);
  ctx->ev.clear();
  ctx->ntasks = 0;

  int64_t max_se = 0, max_nq = 0;
  for (const auto& c : chunks) {
    max_se = std::max(max_se, c.scratch_elems);
    max_nq = std::max(max_nq, c.nq);
  }
  std::vector<double> hx(npts_total), hy(npts_total), hz(npts_total);
  for (int64_t k = 0; k < npts_total; ++k) {
    hx[k] = points[3 * k];
    hy[k] = points[3 * k + 1];
    hz[k] = points[3 * k + 2];
  }
  std::vector<WorkItem> work;
  work.reserve(pbl.size() + g2.size());
  const int64_t base_g2 = (int64_t)pbl.size();
  work.insert(work.end(), pbl.begin(), pbl.end());
  work.insert(work.end(), g2.begin(), g2.end());
  for (auto& c : chunks) c.g2_off += base_g2;

  int rc;
  if ((rc = dalloc(&ctx->d_tasks, ht.size()))) return rc;
  if ((rc = dalloc(&ctx->d_px, npts_total))) return rc;
  if ((rc = dalloc(&ctx->d_py, npts_total))) return rc;
  if ((rc = dalloc(&ctx->d_pz, npts_total))) return rc;
  if ((rc = dalloc(&ctx->d_w, npts_total))) return rc;
  if ((rc = dalloc(&ctx->d_shell_list, (size_t)nshell_entries))) return rc;
  if ((rc = dalloc(&ctx->d_bfmap, bfmap.size()))) return rc;
  if ((rc = dalloc(&ctx->d_lsh, lsh.size()))) return rc;
  if ((rc = dalloc(&ctx->d_work, work.size()))) return rc;
  if ((rc = dalloc(&ctx->d_red, (size_t)2 * std::max<int64_t>(pblocks, 1)))) return rc;
  ctx->scratch_bytes = (size_t)max_se * sizeof(double);
  if (ctx->scratch_bytes) {
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
  XCQ_HIP(h2d(ctx->d_tasks, ht.data(), ht.size() * sizeof(DevTask)));
  XCQ_HIP(h2d(ctx->d_px, hx.data(), npts_total * 8));
  XCQ_HIP(h2d(ctx->d_py, hy.data(), npts_total * 8));
  XCQ_HIP(h2d(ctx->d_pz, hz.data(), npts_total * 8));
  XCQ_HIP(h2d(ctx->d_w, weights, npts_total * 8));
  XCQ_HIP(h2d(ctx->d_shell_list, shell_lists, nshell_entries * 4));
  XCQ_HIP(h2d(ctx->d_bfmap, bfmap.data(), bfmap.size() * 4));
  XCQ_HIP(h2d(ctx->d_lsh, lsh.data(), lsh.size() * 4));
  XCQ_HIP(h2d(ctx->d_work, work.data(), work.size() * sizeof(WorkItem)));

  ctx->ev.resize(chunks.size() * kStages * 2);
  for (auto& e : ctx->ev) XCQ_HIP(hipEventCreate(&e));

  ctx->h_tasks = std::move(ht);
  ctx->chunks = std::move(chunks);
  ctx->ntasks = ntasks;
  ctx->npts_total = npts_total;
  ctx->total_pblocks = pblocks;
  ctx->max_nq = max_nq;
  ctx->max_scratch_elems = max_se;
  int64_t sum_pts = 0, sum_pnn = 0;
  for (const auto& d : ctx->h_tasks) {
    sum_pts += d.npts;
    sum_pnn += (int64_t)d.npts * d.nbe * d.nbe;
  }
  ctx->counters[2] = sum_pts;
  ctx->counters[3] = sum_pnn;
  return XCQ_SUCCESS;
}

int xcq_upload_density(xcq_context ctx, int32_t nbf, const double* P, int64_t ldp) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(P, XCQ_ERROR_INVALID_ARGUMENT, "null density");
  XCQ_REQUIRE(nbf == ctx->nbf && nbf > 0, XCQ_ERROR_INVALID_ARGUMENT,
              "density dimension does not match the basis");
  XCQ_REQUIRE(ldp >= nbf, XCQ_ERROR_INVALID_ARGUMENT, "ldp < nbf");
  XCQ_HIP(hipSetDevice(ctx->device));
  XCQ_HIP(hipMemcpy2DAsync(ctx->d_P, (size_t)nbf * 8, P, (size_t)ldp * 8, (size_t)nbf * 8, nbf,
                           hipMemcpyHostToDevice, ctx->stream));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  ctx->have_density = true;
  if (ctx->d_Phi) {  // invalidate the split copy of the previous density
    (void)hipFree(ctx->d_Phi);
    ctx->d_Phi = nullptr;
  }
  return XCQ_SUCCESS;
}

static int run_pipeline(xcq_context ctx, int func, int precision, int want_vxc) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_REQUIRE(ctx->ntasks > 0, XCQ_ERROR_STATE, "no tasks uploaded");
  XCQ_REQUIRE(ctx->have_density, XCQ_ERROR_STATE, "no density uploaded");
  XCQ_REQUIRE(func == 0 || func == XCQ_FUNC_SVWN5 || func == XCQ_FUNC_PBE || func == XCQ_FUNC_PBE0,
              XCQ_ERROR_INVALID_ARGUMENT, "unknown functional");
  XCQ_REQUIRE(precision == XCQ_PREC_FP64 || precision == XCQ_PREC_MIXED,
              XCQ_ERROR_INVALID_ARGUMENT, "unknown precision");
  XCQ_HIP(hipSetDevice(ctx->device));
  const int gga = (func == XCQ_FUNC_PBE || func == XCQ_FUNC_PBE0) ? 1 : 0;
  hipStream_t s = ctx->stream;
  if (want_vxc)
    XCQ_HIP(hipMemsetAsync(ctx->d_vxc, 0, (size_t)ctx->nbf * ctx->nbf * 8, s));
  int64_t launches = 0, tiles = 0;
  for (size_t ci = 0; ci < ctx->chunks.size(); ++ci) {
    const Chunk& c = ctx->chunks[ci];
    hipEvent_t* ev = ctx->ev.data() + ci * kStages * 2;
    int rc;
    if (precision == XCQ_PREC_MIXED) {
      rc = launch_mixed_pipeline(ctx, c, func, gga, want_vxc, ev);
      if (rc) return rc;
      launches += want_vxc ? 4 : 3;
      tiles += c.npblk + (want_vxc ? c.ng2 : 0);
      continue;
    }
    double* scr = static_cast<double*>(ctx->d_scratch);
    XCQ_HIP(hipEventRecord(ev[0], s));
    rc = launch_collocation_fp64(s, c, ctx->d_work, ctx->d_tasks, ctx->d_shells,
                                 ctx->d_shell_list, ctx->d_lsh, ctx->d_px, ctx->d_py, ctx->d_pz,
                                 scr, gga);
    if (rc) return rc;
    XCQ_HIP(hipEventRecord(ev[1], s));
    XCQ_HIP(hipEventRecord(ev[2], s));
    // X, rho/grad rho, functional, EXC/NEL and Z in one launch (stage 1);
    // stages 2 (functional) and 3 (Z) stay as empty intervals so the stage
    // table keeps its shape.
    rc = launch_density_fp64(s, c, ctx->d_work, ctx->d_tasks, ctx->d_bfmap, ctx->d_P, ctx->nbf,
                             scr, ctx->d_w, func, gga, func == 0 ? 1.0 : 2.0, want_vxc,
                             ctx->d_red);
    if (rc) return rc;
    XCQ_HIP(hipEventRecord(ev[3], s));
    XCQ_HIP(hipEventRecord(ev[4], s));
    XCQ_HIP(hipEventRecord(ev[5], s));
    launches += 2;
    tiles += c.npblk;
    if (want_vxc) {
      XCQ_HIP(hipEventRecord(ev[6], s));
      XCQ_HIP(hipEventRecord(ev[7], s));
      XCQ_HIP(hipEventRecord(ev[8], s));
      rc = launch_syr2k_scatter_fp64(s, c, ctx->d_work, ctx->d_tasks, ctx->d_bfmap, scr,
                                     ctx->d_vxc, ctx->nbf);
      if (rc) return rc;
      XCQ_HIP(hipEventRecord(ev[9], s));
      launches += 1;
      tiles += c.ng2;
    }
  }
  hipEvent_t* evs = ctx->ev.data();  // reuse chunk 0 slots 10/11 for symmetrise
  int rc = 0;
  if (want_vxc) {
    XCQ_HIP(hipEventRecord(evs[10], s));
    rc = launch_symmetrize(s, ctx->nbf, ctx->d_vxc, ctx->nbf);
    if (rc) return rc;
    XCQ_HIP(hipEventRecord(evs[11], s));
    ++launches;
  }
  rc = launch_reduce_sums(s, ctx->d_red, ctx->total_pblocks, ctx->d_out2);
  if (rc) return rc;
  ctx->counters[0] = launches + 1;
  ctx->counters[1] = tiles;
  ctx->last_want_vxc = want_vxc;
  return XCQ_SUCCESS;
}

int xcq_exc_vxc(xcq_context ctx, int functional, int precision, int want_vxc) {
  XCQ_REQUIRE(functional != 0, XCQ_ERROR_INVALID_ARGUMENT, "unknown functional");
  return run_pipeline(ctx, functional, precision, want_vxc ? 1 : 0);
}

int xcq_integrate_den(xcq_context ctx, double* nel) {
  XCQ_REQUIRE(nel, XCQ_ERROR_INVALID_ARGUMENT, "null output");
  int rc = run_pipeline(ctx, 0, XCQ_PREC_FP64, 0);
  if (rc) return rc;
  double out[2];
  XCQ_HIP(hipMemcpyAsync(out, ctx->d_out2, 16, hipMemcpyDeviceToHost, ctx->stream));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  *nel = out[1];
  return XCQ_SUCCESS;
}

int xcq_download_results(xcq_context ctx, double* exc, double* nel, double* vxc, int64_t ldvxc) {
  XCQ_REQUIRE(ctx, XCQ_ERROR_INVALID_ARGUMENT, "null context");
  XCQ_HIP(hipSetDevice(ctx->device));
  double out[2];
  XCQ_HIP(hipMemcpyAsync(out, ctx->d_out2, 16, hipMemcpyDeviceToHost, ctx->stream));
  if (vxc) {
    XCQ_REQUIRE(ctx->last_want_vxc, XCQ_ERROR_STATE, "last evaluation did not compute VXC");
    XCQ_REQUIRE(ldvxc >= ctx->nbf, XCQ_ERROR_INVALID_ARGUMENT, "ldvxc < nbf");
    XCQ_HIP(hipMemcpy2DAsync(vxc, (size_t)ldvxc * 8, ctx->d_vxc, (size_t)ctx->nbf * 8,
                             (size_t)ctx->nbf * 8, ctx->nbf, hipMemcpyDeviceToHost,
                             ctx->stream));
  }
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  if (exc) *exc = out[0];
  if (nel) *nel = out[1];
  return XCQ_SUCCESS;
}

int xcq_stage_times(xcq_context ctx, double* ms, int n) {
  XCQ_REQUIRE(ctx && ms, XCQ_ERROR_INVALID_ARGUMENT, "null argument");
  XCQ_HIP(hipSetDevice(ctx->device));
  XCQ_HIP(hipStreamSynchronize(ctx->stream));
  double st[kStages + 1] = {0};
  const int nst = ctx->last_want_vxc ? 5 : 3;
  for (size_t ci = 0; ci < ctx->chunks.size(); ++ci)
    for (int k = 0; k < nst; ++k) {
      float f = 0.f;
      hipEvent_t* ev = ctx->ev.data() + ci * kStages * 2;
      if (hipEventElapsedTime(&f, ev[2 * k], ev[2 * k + 1]) == hipSuccess) st[k] += f;
      else (void)hipGetLastError();
    }
  if (ctx->last_want_vxc && !ctx->ev.empty()) {
    float f = 0.f;
    if (hipEventElapsedTime(&f, ctx->ev[10], ctx->ev[11]) == hipSuccess) st[5] = f;
    else (void)hipGetLastError();
  }
  for (int k = 0; k < kStages; ++k) st[kStages] += st[k];
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
  // Inter-atomic distances and nearest-neighbour distances (MolMeta).
  std::vector<double> rab((size_t)natoms * natoms, 0.), dn(natoms, INFINITY);
  for (int i = 0; i < natoms; ++i)
    for (int j = 0; j < natoms; ++j) {
      const double dx = atom_xyz[3 * i] - atom_xyz[3 * j];
      const double dy = atom_xyz[3 * i + 1] - atom_xyz[3 * j + 1];
      const double dz = atom_xyz[3 * i + 2] - atom_xyz[3 * j + 2];
      rab[(size_t)i + (size_t)j * natoms] = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (i != j) dn[i] = std::min(dn[i], rab[(size_t)i + (size_t)j * natoms]);
    }
  std::vector<double> hx(npts), hy(npts), hz(npts);
  for (int64_t k = 0; k < npts; ++k) {
    hx[k] = points[3 * k];
    hy[k] = points[3 * k + 1];
    hz[k] = points[3 * k + 2];
  }
  double *d_xyz = nullptr, *d_rab = nullptr, *d_dn = nullptr, *d_x = nullptr, *d_y = nullptr,
         *d_z = nullptr, *d_w = nullptr;
  int32_t* d_par = nullptr;
  int rc = XCQ_SUCCESS;
  auto cleanup = [&] {
    dfree(d_xyz);
    dfree(d_rab);
    dfree(d_dn);
    dfree(d_x);
    dfree(d_y);
    dfree(d_z);
    dfree(d_w);
    dfree(d_par);
  };
  if ((rc = dalloc(&d_xyz, 3 * (size_t)natoms)) || (rc = dalloc(&d_rab, rab.size())) ||
      (rc = dalloc(&d_dn, (size_t)natoms)) || (rc = dalloc(&d_x, (size_t)npts)) ||
      (rc = dalloc(&d_y, (size_t)npts)) || (rc = dalloc(&d_z, (size_t)npts)) ||
      (rc = dalloc(&d_w, (size_t)npts)) || (rc = dalloc(&d_par, (size_t)npts))) {
    cleanup();
    return rc;
  }
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMemcpy(d_xyz, atom_xyz, 24 * (size_t)natoms, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_rab, rab.data(), 8 * rab.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_dn, dn.data(), 8 * (size_t)natoms, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_x, hx.data(), 8 * (size_t)npts, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_y, hy.data(), 8 * (size_t)npts, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_z, hz.data(), 8 * (size_t)npts, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_w, weights, 8 * (size_t)npts, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_par, parent, 4 * (size_t)npts, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error(std::string("SSF upload: ") + hipGetErrorString(e));
    cleanup();
    return XCQ_ERROR_HIP;
  }
  rc = launch_ssf(nullptr, natoms, d_xyz, d_rab, d_dn, npts, d_x, d_y, d_z, d_par, d_w);
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

int xcq_ssf_weights_tasks(xcq_context ctx, int32_t natoms, const double* atom_xyz) {
  XCQ_REQUIRE(ctx && ctx->ntasks > 0, XCQ_ERROR_STATE, "no tasks uploaded");
  XCQ_REQUIRE(natoms > 0 && atom_xyz, XCQ_ERROR_INVALID_ARGUMENT, "empty molecule");
  XCQ_HIP(hipSetDevice(ctx->device));
  std::vector<int32_t> parent(ctx->npts_total);
  for (const auto& d : ctx->h_tasks) {
    XCQ_REQUIRE(d.iparent >= 0 && d.iparent < natoms, XCQ_ERROR_INVALID_ARGUMENT,
                "task parent atom out of range");
    for (int32_t i = 0; i < d.npts; ++i) parent[d.pt_off + i] = d.iparent;
  }
  std::vector<double> rab((size_t)natoms * natoms, 0.), dn(natoms, INFINITY);
  for (int i = 0; i < natoms; ++i)
    for (int j = 0; j < natoms; ++j) {
      const double dx = atom_xyz[3 * i] - atom_xyz[3 * j];
      const double dy = atom_xyz[3 * i + 1] - atom_xyz[3 * j + 1];
      const double dz = atom_xyz[3 * i + 2] - atom_xyz[3 * j + 2];
      rab[(size_t)i + (size_t)j * natoms] = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (i != j) dn[i] = std::min(dn[i], rab[(size_t)i + (size_t)j * natoms]);
    }
  double *d_xyz = nullptr, *d_rab = nullptr, *d_dn = nullptr;
  int32_t* d_par = nullptr;
  int rc;
  if ((rc = dalloc(&d_xyz, 3 * (size_t)natoms)) || (rc = dalloc(&d_rab, rab.size())) ||
      (rc = dalloc(&d_dn, (size_t)natoms)) || (rc = dalloc(&d_par, parent.size()))) {
    dfree(d_xyz);
    dfree(d_rab);
    dfree(d_dn);
    dfree(d_par);
    return rc;
  }
  hipError_t e = hipMemcpy(d_xyz, atom_xyz, 24 * (size_t)natoms, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_rab, rab.data(), 8 * rab.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_dn, dn.data(), 8 * (size_t)natoms, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(d_par, parent.data(), 4 * parent.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    rc = launch_ssf(ctx->stream, natoms, d_xyz, d_rab, d_dn, ctx->npts_total, ctx->d_px,
                    ctx->d_py, ctx->d_pz, d_par, ctx->d_w);
    if (rc == XCQ_SUCCESS) e = hipStreamSynchronize(ctx->stream);
  }
  dfree(d_xyz);
  dfree(d_rab);
  dfree(d_dn);
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
  const DevTask& d = ctx->h_tasks[task];
  const int nmb = d.npts_pad / kPtBlock;
  std::vector<WorkItem> items(nmb);
  for (int pb = 0; pb < nmb; ++pb) items[pb] = WorkItem{task, pb, 0, 0};
  WorkItem* d_items = nullptr;
  int rc = dalloc(&d_items, items.size());
  if (rc) return rc;
  Chunk c;
  std::memset(&c, 0, sizeof(c));
  c.npblk = nmb;
  const int ncomp = deriv ? 4 : 1;
  const int64_t plane = (int64_t)d.nbe * d.npts_pad;
  std::vector<double> host((size_t)ncomp * plane);
  hipError_t e = hipMemcpy(d_items, items.data(), items.size() * sizeof(WorkItem),
                           hipMemcpyHostToDevice);
  if (e == hipSuccess && nmb > 0) {
    if (precision == XCQ_PREC_FP64) {
      double* scr = static_cast<double*>(ctx->d_scratch);
      rc = launch_collocation_fp64(ctx->stream, c, d_items, ctx->d_tasks, ctx->d_shells,
                                   ctx->d_shell_list, ctx->d_lsh, ctx->d_px, ctx->d_py,
                                   ctx->d_pz, scr, deriv);
      if (rc == XCQ_SUCCESS) e = hipStreamSynchronize(ctx->stream);
      if (rc == XCQ_SUCCESS && e == hipSuccess)
        e = hipMemcpy(host.data(), scr + d.b_off, 8 * host.size(), hipMemcpyDeviceToHost);
    } else {
      // FP32 collocation: same layout in float inside the scratch.
      float* scr = static_cast<float*>(ctx->d_scratch);
      rc = launch_collocation_fp32(ctx->stream, c, d_items, ctx->d_tasks, ctx->d_shells,
                                   ctx->d_shell_list, ctx->d_lsh, ctx->d_px, ctx->d_py,
                                   ctx->d_pz, scr, deriv);
      if (rc == XCQ_SUCCESS) e = hipStreamSynchronize(ctx->stream);
      std::vector<float> hf(host.size());
      if (rc == XCQ_SUCCESS && e == hipSuccess)
        e = hipMemcpy(hf.data(), scr + d.b_off, 4 * hf.size(), hipMemcpyDeviceToHost);
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
