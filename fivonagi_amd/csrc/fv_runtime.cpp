// fv_runtime.cpp -- host side of the C ABI (include/fivonagi.h): device
// memory, instance tables, work lists, hipGraph replay of step launches.
//
// The reference's host loop (common/main.cu:310-635) copies dt and simError
// to the host up to seven times per step; here the step kernel keeps dt, T,
// the accepted/rejected counters and the buffer index on the device, and the
// host only synchronises once per fivonagi_step call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/fivonagi.h"
#include "fv_adjoint.h"
#include "fv_tables.h"
#include "fv_types.h"

extern "C" hipError_t fv_launch_step(const fv::InstDesc* insts, const fv::WorkItem* items,
                                     const uint64_t* rowmask, int n_items, int phase,
                                     int n_inst, fv::Ctl* ctl0, fv::SyncSlots* sync0, int bc0,
                                     int ext, int fp64, hipStream_t stream);
extern "C" hipError_t fv_launch_settle(const fv::InstDesc* insts, int n_inst, int phase,
                                       hipStream_t stream);
extern "C" hipError_t fv_launch_collect(const fv::InstDesc* insts, int n_inst, int max_nx,
                                        int phase, int fp64, hipStream_t stream);
extern "C" hipError_t fv_step_waves_per_cu(int* waves, int fp64);
extern "C" hipError_t fv_adj_reverse(const fv::AdjDesc* ds, const fv::AdjItem* items,
                                     const int32_t* wave_first, int n_waves, int k, int slot,
                                     int par, int check, hipStream_t s);
extern "C" hipError_t fv_adj_occupancy(int* waves_per_cu);
extern "C" hipError_t fv_launch_taylor_ghosts(const fv::InstDesc* insts, int n_inst,
                                              int max_ring, int phase, hipStream_t stream);
extern "C" hipError_t fv_scratch_get(void** ptr, size_t* bytes);

using fvh::auto_rows;
using fvh::fold_subset;
using fvh::HostInst;
using fvh::host_inst;
using fvh::table_bytes;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return -(int)e;
}

#define HIP_TRY(expr)                               \
  do {                                              \
    hipError_t e_ = (expr);                         \
    if (e_ != hipSuccess) return hip_fail(e_, #expr); \
  } while (0)

}  // namespace

struct fivonagi_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  int n_inst = 0;
  std::vector<fivonagi_params> params;
  std::vector<fv::InstDesc> desc;          // host copy of the instance table
  std::vector<fv::WorkItem> items;
  std::vector<uint64_t> rowmask;           // CFL lane mask per item row (fv::WorkItem::mrow)
  bool fp64 = false;                       // state and arithmetic in double
  bool md_any = false;                     // some instance has a moving window
  bool ext_any = false;                    // ... or the HIFU boundary (extended kernel)
  float skew = -1.0f;                      // height skew of the work items (< 0: automatic)
  float skew_used = 0.0f;                  // the skew of the current work list
  bool rotate = false;                     // level pattern rotated per strip (chunk_bounds)
  bool rotate_used = false;
  long cus = 256;                          // CUs of the device
  int per_cu = 16;                         // resident step waves per CU
  std::vector<std::vector<uint32_t>> fold_bits;  // kCflBitmap: bit i + j*nx = cell counted
  // the caller's edge_amp / edge_phase arrays, copied at init (params point here)
  std::vector<std::vector<double>> edge_amp, edge_phase;
  // prescribed dt sequence per instance: device buffer and its capacity
  std::vector<double*> dt_seq_dev;
  std::vector<int64_t> dt_seq_cap;
  std::vector<void*> allocs;               // every device allocation
  fv::InstDesc* d_desc = nullptr;
  fv::WorkItem* d_items = nullptr;
  uint64_t* d_rowmask = nullptr;
  fv::Ctl* d_ctl = nullptr;                // 2 control blocks per instance
  fv::SyncSlots* d_sync = nullptr;         // 3 reduction slot sets per instance
  int64_t attempts = 0;                    // step attempts launched (the phase is this mod 6)
  // captured step launches, each starting at phase 0: requests of up to
  // kGraphChunk attempts replay one graph of their exact length, longer ones
  // graphs of kGraphChunk (a multiple of 6, so it ends at phase 0) and one of
  // the rest; at most kGraphCache lengths are kept (the oldest goes first)
  static constexpr int kGraphChunk = 300, kGraphCache = 8;
  int taylor_ring = 0;                     // largest ghost ring of a Taylor instance
  int collect_nx = 0;                      // widest instance with data collection (0: none)
  bool keep_hist = false;                  // the step stores accepted states into the
                                           // instances' histories (adjoint forward pass)
  bool unlimited = false;                  // targets set to "run every attempt"
  struct Graph {
    int len;
    hipGraph_t g;
    hipGraphExec_t e;
  };
  std::vector<Graph> graphs;                 // oldest first
  std::vector<float*> adj_lam;             // dJ/dq0 of the last adjoint call (per instance)
  void* adj_pool = nullptr;                // the adjoint calls' device buffers, kept for
  size_t adj_pool_bytes = 0;               // the next call (grown when it needs more)
  struct CollectBufs { void* measure = nullptr; void* stage = nullptr; };
  std::vector<CollectBufs> collect_bufs;   // per instance, kept once allocated
  uint32_t* d_collect_err = nullptr;       // collection overflow flag per instance
  // control blocks as the host last read them; valid until the next launch
  // or host write (read_ctls, launch_attempts, write_ctls)
  std::vector<fv::Ctl> ctl_cache;
  bool ctl_cache_ok = false;

  void* dalloc(size_t bytes, hipError_t* e) {
    void* p = nullptr;
    *e = hipMalloc(&p, bytes);
    if (*e == hipSuccess) allocs.push_back(p);
    return p;
  }
  void drop_graph() {
    for (auto& gr : graphs) {
      (void)hipGraphExecDestroy(gr.e);
      (void)hipGraphDestroy(gr.g);
    }
    graphs.clear();
  }
  ~fivonagi_ctx() {
    drop_graph();
    if (stream) (void)hipStreamSynchronize(stream);
    for (float* q : adj_lam) (void)hipFree(q);
    if (adj_pool) (void)hipFree(adj_pool);
    for (double* d : dt_seq_dev) (void)hipFree(d);
    if (d_items) (void)hipFree(d_items);
    if (d_rowmask) (void)hipFree(d_rowmask);
    for (void* p : allocs) (void)hipFree(p);
    if (own_stream) (void)hipStreamDestroy(own_stream);
  }
};

namespace {

// CUs of the device and resident step waves per CU (4 SIMDs per CU)
void device_shape(int device, bool fp64, long* cus, int* per_cu) {
  int n = 0, w = 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      fv_step_waves_per_cu(&w, fp64 ? 1 : 0) != hipSuccess || n <= 0 || w <= 0) {
    n = 256;   // MI355X: 256 CUs, 4 waves of the FP32 step kernel per SIMD, 2 of FP64
    w = fp64 ? 8 : 16;
  }
  *cus = n;
  *per_cu = w;
}

// The step kernel's work list for the context's instances (fv_tables.cpp).
void build_items(fivonagi_ctx* c, int auto_h, float skew, bool rotate) {
  std::vector<int> masks(c->n_inst);
  for (int k = 0; k < c->n_inst; ++k) masks[k] = c->desc[k].cfl_mask;
  fvh::WorkList wl;
  fvh::build_work(c->params, masks, c->fold_bits, c->cus, c->per_cu, auto_h, skew, rotate, wl);
  c->items = std::move(wl.items);
  c->rowmask = std::move(wl.rowmask);
  for (int k = 0; k < c->n_inst; ++k) c->desc[k].n_items = wl.n_items[k];
  c->skew_used = wl.skew_used;
  c->rotate_used = wl.rotate_used;
}

// (re)upload the work list after build_items; the captured graph holds the
// old grid size, so it is dropped
int upload_items(fivonagi_ctx* c) {
  c->drop_graph();
  c->ctl_cache_ok = false;
  if (c->d_items) (void)hipFree(c->d_items);
  c->d_items = nullptr;
  if (c->d_rowmask) (void)hipFree(c->d_rowmask);
  c->d_rowmask = nullptr;
  HIP_TRY(hipMalloc(&c->d_items, sizeof(fv::WorkItem) * c->items.size()));
  HIP_TRY(hipMemcpyAsync(c->d_items, c->items.data(), sizeof(fv::WorkItem) * c->items.size(),
                         hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMalloc(&c->d_rowmask, sizeof(uint64_t) * c->rowmask.size()));
  HIP_TRY(hipMemcpyAsync(c->d_rowmask, c->rowmask.data(), sizeof(uint64_t) * c->rowmask.size(),
                         hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->d_desc, c->desc.data(), sizeof(fv::InstDesc) * c->n_inst,
                         hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int upload_tables(fivonagi_ctx* c) {
  c->ctl_cache_ok = false;   // the settled control blocks depend on the descriptors
  HIP_TRY(hipMemcpyAsync(c->d_desc, c->desc.data(), sizeof(fv::InstDesc) * c->n_inst,
                         hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

// dt-derived step scalars, as the step kernel's finalizer forms them
void set_ctl_derived(fv::Ctl& t, const fv::InstDesc& d) {
  t.dtx = (float)(t.dt * d.inv_dx);
  t.dty = (float)(t.dt * d.inv_dy);
  t.ddx2 = (float)(t.dt * d.dif_x);
  t.ddy2 = (float)(t.dt * d.dif_y);
  const double turns = d.omega_turns * t.T;
  const double th = 6.283185307179586476925286766559 * (turns - std::floor(turns));
  t.drive_s = std::sin(th);
  t.drive_c = std::cos(th);
}

// The control blocks after the last launched attempt: the settle kernel
// folds that attempt's reduction into block (a+1)%2 (a = next attempt index),
// which is where the next launch will store the same result.
int read_ctls(fivonagi_ctx* c, std::vector<fv::Ctl>& out) {
  if (c->ctl_cache_ok) {
    out = c->ctl_cache;
    return 0;
  }
  const int phase = (int)(c->attempts % 6);
  HIP_TRY(fv_launch_settle(c->d_desc, c->n_inst, phase, c->stream));
  std::vector<fv::Ctl> both(2 * (size_t)c->n_inst);
  HIP_TRY(hipMemcpyAsync(both.data(), c->d_ctl, sizeof(fv::Ctl) * both.size(),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  out.resize(c->n_inst);
  const int b = (phase + 1) & 1;
  for (int k = 0; k < c->n_inst; ++k) out[k] = both[2 * (size_t)k + b];
  c->ctl_cache = out;
  c->ctl_cache_ok = true;
  return 0;
}

// Replace the control blocks the next attempt starts from: block a%2, with
// the reduction of attempt a-1 marked as already folded in.
int write_ctls(fivonagi_ctx* c, const std::vector<fv::Ctl>& in) {
  c->ctl_cache_ok = false;
  const int phase = (int)(c->attempts % 6);
  const uint32_t zero = 0u;
  for (int k = 0; k < c->n_inst; ++k) {
    HIP_TRY(hipMemcpyAsync(c->d_ctl + 2 * (size_t)k + (phase & 1), &in[k], sizeof(fv::Ctl),
                           hipMemcpyHostToDevice, c->stream));
    fv::SyncSlots* prev = c->d_sync + 3 * (size_t)k + (phase + 2) % 3;
    HIP_TRY(hipMemcpyAsync(&prev->ran, &zero, sizeof(uint32_t), hipMemcpyHostToDevice,
                           c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

// One step attempt: the Taylor ghost pre-pass (only if a Taylor instance
// exists) and the fused step, for attempt index = phase (mod 6).
hipError_t launch_phase(fivonagi_ctx* c, int phase) {
  if (c->taylor_ring > 0) {
    hipError_t e =
        fv_launch_taylor_ghosts(c->d_desc, c->n_inst, c->taylor_ring, phase, c->stream);
    if (e != hipSuccess) return e;
  }
  hipError_t e = fv_launch_step(c->d_desc, c->d_items, c->d_rowmask, (int)c->items.size(),
                                phase, c->n_inst, c->d_ctl, c->d_sync, c->params[0].bc,
                                c->ext_any || c->keep_hist ? 1 : 0, c->fp64 ? 1 : 0,
                                c->stream);
  if (e == hipSuccess && c->collect_nx > 0)
    e = fv_launch_collect(c->d_desc, c->n_inst, c->collect_nx, phase, c->fp64 ? 1 : 0,
                          c->stream);
  return e;
}

// data collection overflow flags (any instance collecting): FIVONAGI_ERR_COLLECT or 0
int collect_status(fivonagi_ctx* c) {
  if (!c->d_collect_err) return 0;
  std::vector<uint32_t> v(c->n_inst);
  HIP_TRY(hipMemcpyAsync(v.data(), c->d_collect_err, sizeof(uint32_t) * c->n_inst,
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int k = 0; k < c->n_inst; ++k)
    if (v[k] && c->desc[k].measure)
      return fail(FIVONAGI_ERR_COLLECT, "data collection: more than nx steps in a frame");
  return 0;
}

// The captured graph of G attempts from phase 0 (phases 0..G-1 mod 6),
// captured on first use.
int graph_for(fivonagi_ctx* c, int G, hipGraphExec_t* out) {
  for (auto& gr : c->graphs)
    if (gr.len == G) {
      *out = gr.e;
      return 0;
    }
  if ((int)c->graphs.size() >= fivonagi_ctx::kGraphCache) {
    (void)hipGraphExecDestroy(c->graphs.front().e);
    (void)hipGraphDestroy(c->graphs.front().g);
    c->graphs.erase(c->graphs.begin());
  }
  HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  for (int a = 0; a < G; ++a) {
    hipError_t e = launch_phase(c, a % 6);
    if (e != hipSuccess) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(c->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      return hip_fail(e, "fv_launch_step (capture)");
    }
  }
  fivonagi_ctx::Graph gr{G, nullptr, nullptr};
  HIP_TRY(hipStreamEndCapture(c->stream, &gr.g));
  hipError_t e = hipGraphInstantiate(&gr.e, gr.g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(gr.g);
    return hip_fail(e, "hipGraphInstantiate");
  }
  c->graphs.push_back(gr);
  *out = gr.e;
  return 0;
}

// n attempts: single launches up to a phase-0 boundary, then graph replays
// (graph_for: the exact length of a request up to kGraphChunk, else chunks
// and the rest); fewer than 6 attempts run as single launches.
int launch_attempts(fivonagi_ctx* c, int64_t n) {
  c->ctl_cache_ok = false;
  while (n > 0 && c->attempts % 6 != 0) {
    hipError_t e = launch_phase(c, (int)(c->attempts % 6));
    if (e != hipSuccess) return hip_fail(e, "fv_launch_step");
    ++c->attempts;
    --n;
  }
  while (n >= 6) {
    const int G = (int)std::min<int64_t>(n, fivonagi_ctx::kGraphChunk);
    hipGraphExec_t ge = nullptr;
    if (int rc = graph_for(c, G, &ge)) return rc;
    HIP_TRY(hipGraphLaunch(ge, c->stream));
    c->attempts += G;
    n -= G;
    // a graph of a length that is not a multiple of 6 ends mid-rotation
    while (n > 0 && c->attempts % 6 != 0) {
      hipError_t e = launch_phase(c, (int)(c->attempts % 6));
      if (e != hipSuccess) return hip_fail(e, "fv_launch_step");
      ++c->attempts;
      --n;
    }
  }
  while (n > 0) {
    hipError_t e = launch_phase(c, (int)(c->attempts % 6));
    if (e != hipSuccess) return hip_fail(e, "fv_launch_step");
    ++c->attempts;
    --n;
  }
  return 0;
}

int check_ctx(fivonagi_ctx* c, int inst) {
  if (!c) return fail(FIVONAGI_ERR_ARG, "null context");
  if (inst < 0 || inst >= c->n_inst) return fail(FIVONAGI_ERR_ARG, "instance out of range");
  return 0;
}

// Stop conditions of the attempts launched from now on: an instance runs
// while n_acc < tgt[k] and T < T_stop.
int write_targets(fivonagi_ctx* c, const std::vector<int64_t>& tgt,
                  double T_stop = std::numeric_limits<double>::infinity()) {
  std::vector<fv::Ctl> ctl;
  if (int rc = read_ctls(c, ctl)) return rc;
  for (int k = 0; k < c->n_inst; ++k) {
    ctl[k].target_acc = tgt[k];
    ctl[k].T_stop = T_stop;
  }
  return write_ctls(c, ctl);
}

// Parameters every entry point that sets up an instance accepts.
int check_params(const fivonagi_params& q, bool f64) {
  std::string why;
  const int rc = fvh::check_params(q, f64, &why);
  return rc ? fail(rc, why) : 0;
}

// Control block of an instance's first attempt: dt0 at T0, no counts, no
// stop conditions yet.
void initial_ctl(const fivonagi_params& q, const fv::InstDesc& d, fv::Ctl& t) {
  std::memset(&t, 0, sizeof(t));
  t.dt = q.dt0;
  t.T = q.T0;
  t.target_acc = 0;
  t.T_stop = std::numeric_limits<double>::infinity();
  set_ctl_derived(t, d);
}

}  // namespace

extern "C" {

const char* fivonagi_last_error(void) { return g_err.c_str(); }

// fivonagi_init / fivonagi_init_f64: q_host holds float or double grids.
static int init_impl(fivonagi_ctx** out, const fivonagi_params* p, int n_inst,
                     const void* const* q_host, bool f64) {
  if (!out || !p || !q_host || n_inst <= 0) return fail(FIVONAGI_ERR_ARG, "null argument");
  *out = nullptr;
  for (int k = 0; k < n_inst; ++k) {
    if (int rc = check_params(p[k], f64)) return rc;
    if (!q_host[3 * k] || !q_host[3 * k + 1] || !q_host[3 * k + 2])
      return fail(FIVONAGI_ERR_ARG, "null state pointer");
  }
  auto* c = new fivonagi_ctx();
  c->n_inst = n_inst;
  c->fp64 = f64;
  const size_t es = f64 ? sizeof(double) : sizeof(float);
  c->params.assign(p, p + n_inst);
  c->desc.resize(n_inst);
  c->dt_seq_dev.assign(n_inst, nullptr);
  c->dt_seq_cap.assign(n_inst, 0);
  c->fold_bits.resize(n_inst);
  // the edge arrays are copied: the caller may free or reuse its own after
  // init, and the adjoint reads these copies (the drive the forward runs)
  c->edge_amp.resize(n_inst);
  c->edge_phase.resize(n_inst);
  for (int k = 0; k < n_inst; ++k) {
    auto& q = c->params[k];
    if (q.edge_amp) {
      c->edge_amp[k].assign(q.edge_amp, q.edge_amp + q.ny);
      q.edge_amp = c->edge_amp[k].data();
    }
    if (q.edge_phase) {
      c->edge_phase[k].assign(q.edge_phase, q.edge_phase + q.ny);
      q.edge_phase = c->edge_phase[k].data();
    }
  }
  p = c->params.data();
  hipError_t e;
  e = hipGetDevice(&c->device);
  if (e != hipSuccess) { delete c; return hip_fail(e, "hipGetDevice"); }
  e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete c; return hip_fail(e, "hipStreamCreate"); }
  c->stream = c->own_stream;

  c->d_ctl = (fv::Ctl*)c->dalloc(sizeof(fv::Ctl) * 2 * n_inst, &e);
  if (e != hipSuccess) { delete c; return hip_fail(e, "hipMalloc ctl"); }
  c->d_sync = (fv::SyncSlots*)c->dalloc(sizeof(fv::SyncSlots) * 3 * n_inst, &e);
  if (e != hipSuccess) { delete c; return hip_fail(e, "hipMalloc sync"); }
  e = hipMemset(c->d_sync, 0, sizeof(fv::SyncSlots) * 3 * n_inst);
  if (e != hipSuccess) { delete c; return hip_fail(e, "hipMemset sync"); }
  std::vector<fv::Ctl> ctl(2 * (size_t)n_inst);
  for (int k = 0; k < n_inst; ++k) {
    const auto& q = p[k];
    HostInst h;
    host_inst(q, f64, h);
    auto& d = c->desc[k];
    d = h.d;
    const size_t n = (size_t)q.nx * q.ny;
    char* buf = (char*)c->dalloc(es * n * 6, &e);
    if (e != hipSuccess) { delete c; return hip_fail(e, "hipMalloc state"); }
    for (int b = 0; b < 2; ++b)
      for (int f = 0; f < 3; ++f) d.q[b][f] = (float*)(buf + es * (size_t)(3 * b + f) * n);
    for (int f = 0; f < 3; ++f) {
      e = hipMemcpy(d.q[0][f], q_host[3 * k + f], es * n, hipMemcpyHostToDevice);
      if (e != hipSuccess) { delete c; return hip_fail(e, "hipMemcpy state"); }
      e = hipMemcpy(d.q[1][f], q_host[3 * k + f], es * n, hipMemcpyHostToDevice);
      if (e != hipSuccess) { delete c; return hip_fail(e, "hipMemcpy state"); }
    }
    // a table in the context's precision on the device
    auto upload = [&](const std::vector<double>& v, const void** dst) -> hipError_t {
      hipError_t r;
      void* t = c->dalloc(es * v.size(), &r);
      if (r != hipSuccess) return r;
      *dst = t;
      std::vector<char> bytes(es * v.size());
      table_bytes(v, f64, bytes.data());
      return hipMemcpy(t, bytes.data(), bytes.size(), hipMemcpyHostToDevice);
    };
    if (!h.tab.empty()) {
      e = upload(h.tab, &d.hifu_tab);
      if (e != hipSuccess) { delete c; return hip_fail(e, "hifu table"); }
    }
    e = upload(h.ec, &d.edge_c);
    if (e == hipSuccess) e = upload(h.es, &d.edge_s);
    if (e != hipSuccess) { delete c; return hip_fail(e, "edge tables"); }
    c->fold_bits[k] = std::move(h.fold_bits);
    if (h.ext) c->ext_any = true;
    if (q.md_on) c->md_any = true;
    c->taylor_ring = std::max(c->taylor_ring, h.taylor_ring);
    d.ctl = c->d_ctl + 2 * k;
    d.sync = c->d_sync + 3 * k;
    // attempt 0 starts from block 0; slot set 2 (attempt -1) never ran
    initial_ctl(q, d, ctl[2 * (size_t)k]);
    ctl[2 * (size_t)k + 1] = ctl[2 * (size_t)k];
  }
  e = hipMemcpy(c->d_ctl, ctl.data(), sizeof(fv::Ctl) * 2 * n_inst, hipMemcpyHostToDevice);
  if (e != hipSuccess) { delete c; return hip_fail(e, "hipMemcpy ctl"); }

  device_shape(c->device, c->fp64, &c->cus, &c->per_cu);
  build_items(c, auto_rows(c->params, c->cus * c->per_cu), c->skew, c->rotate);
  c->d_desc = (fv::InstDesc*)c->dalloc(sizeof(fv::InstDesc) * n_inst, &e);
  if (e != hipSuccess) { delete c; return hip_fail(e, "hipMalloc desc"); }
  int rc = upload_items(c);
  if (rc) { delete c; return rc; }
  *out = c;
  return 0;
}

int fivonagi_init(fivonagi_ctx** out, const fivonagi_params* p, int n_inst,
                  const float* const* q_host) {
  return init_impl(out, p, n_inst, (const void* const*)q_host, false);
}

int fivonagi_init_f64(fivonagi_ctx** out, const fivonagi_params* p, int n_inst,
                      const double* const* q_host) {
  return init_impl(out, p, n_inst, (const void* const*)q_host, true);
}

int fivonagi_set_stream(fivonagi_ctx* c, void* s) {
  if (!c) return fail(FIVONAGI_ERR_ARG, "null context");
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->drop_graph();
  c->stream = s ? (hipStream_t)s : c->own_stream;
  return 0;
}

int fivonagi_set_dt_sequence(fivonagi_ctx* c, int inst, const double* dt, int64_t n) {
  if (int rc = check_ctx(c, inst)) return rc;
  if (dt && n <= 0) return fail(FIVONAGI_ERR_ARG, "empty dt sequence");
  HIP_TRY(hipStreamSynchronize(c->stream));
  std::vector<fv::Ctl> ctl;
  if (int rc = read_ctls(c, ctl)) return rc;
  auto& d = c->desc[inst];
  if (dt) {
    for (int64_t k = 0; k < n; ++k)
      if (!(dt[k] > 0)) return fail(FIVONAGI_ERR_ARG, "dt sequence values must be > 0");
    // the sequence is indexed by the attempt counter: it must hold an entry
    // for the next attempt (checked before anything changes)
    auto& t = ctl[inst];
    if (t.n_att >= n) return fail(FIVONAGI_ERR_ARG, "dt sequence shorter than attempts made");
    // one device buffer per instance, reused while it is large enough
    if (c->dt_seq_cap[inst] < n) {
      double* dd = nullptr;
      HIP_TRY(hipMalloc(&dd, sizeof(double) * n));
      if (c->dt_seq_dev[inst]) (void)hipFree(c->dt_seq_dev[inst]);
      c->dt_seq_dev[inst] = dd;
      c->dt_seq_cap[inst] = n;
    }
    HIP_TRY(hipMemcpy(c->dt_seq_dev[inst], dt, sizeof(double) * n, hipMemcpyHostToDevice));
    d.dt_seq = c->dt_seq_dev[inst];
    d.dt_seq_len = n;
    t.dt = dt[t.n_att];
    set_ctl_derived(t, d);
    if (int rc = write_ctls(c, ctl)) return rc;
  } else {
    d.dt_seq = nullptr;
    d.dt_seq_len = 0;
  }
  return upload_tables(c);
}

int fivonagi_autotune(fivonagi_ctx* c, int trial_steps, int32_t* rows_out) {
  if (!c || trial_steps <= 0) return fail(FIVONAGI_ERR_ARG, "bad argument");
  std::vector<fv::Ctl> saved;
  if (int rc = read_ctls(c, saved)) return rc;
  for (int k = 0; k < c->n_inst; ++k)
    if (saved[k].err) return fivonagi_sync(c);
  // keep every instance's current state
  std::vector<float*> keep(c->n_inst, nullptr);
  auto release = [&]() {
    for (float* p : keep) (void)hipFree(p);
  };
  for (int k = 0; k < c->n_inst; ++k) {
    const size_t n = (size_t)c->params[k].nx * c->params[k].ny;
    const size_t es = c->fp64 ? sizeof(double) : sizeof(float);
    hipError_t e = hipMalloc(&keep[k], es * 3 * n);
    if (e != hipSuccess) { release(); return hip_fail(e, "hipMalloc (autotune)"); }
    for (int f = 0; f < 3; ++f)
      (void)hipMemcpyAsync((char*)keep[k] + es * (size_t)f * n, c->desc[k].q[saved[k].cur][f],
                           es * n, hipMemcpyDeviceToDevice, c->stream);
  }
  // the trials run with data collection and prescribed dt sequences switched
  // off: they must neither record into the measurement grid (or its overflow
  // flag) nor use up the sequence; both come back with the saved state
  const std::vector<fv::InstDesc> desc_saved = c->desc;
  const int collect_nx_saved = c->collect_nx;
  for (auto& d : c->desc) {
    d.dt_seq = nullptr;
    d.dt_seq_len = 0;
    d.measure = nullptr;
    d.stage = nullptr;
  }
  c->collect_nx = 0;
  const int auto_h = auto_rows(c->params, c->cus * c->per_cu);
  std::vector<int> cands = {auto_h, 8, 12, 16, 20, 24, 32, 40, 48, 64, 96, 128};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  float best_ms = 1e30f;
  int best_h = auto_h;
  float best_skew = c->skew;
  bool best_rot = c->rotate;
  int rc = 0;
  std::vector<fivonagi_params> orig = c->params;
  // one candidate split: trial_steps attempts after 2 warm-up attempts
  auto trial = [&](int h, float skew, bool rot) -> int {
    for (auto& p : c->params) p.rows_per_wave = h;
    build_items(c, h, skew, rot);
    if (int r = upload_items(c)) return r;
    std::vector<int64_t> tgt(c->n_inst, INT64_MAX);
    if (int r = write_targets(c, tgt)) return r;
    if (int r = launch_attempts(c, 2)) return r;
    (void)hipEventRecord(e0, c->stream);
    if (int r = launch_attempts(c, trial_steps)) return r;
    (void)hipEventRecord(e1, c->stream);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms < best_ms) { best_ms = ms; best_h = h; best_skew = skew; best_rot = rot; }
    return 0;
  };
  // heights with the automatic skew, then skews and level layouts at the
  // best height
  for (int h : cands)
    if ((rc = trial(h, -1.0f, false))) break;
  if (!rc) {
    const int h = best_h;
    for (float sk : {0.0f, 0.25f, 0.5f, 0.75f})
      for (bool rot : {false, true}) {
        if (sk == 0.0f && rot) continue;
        if (!rc) rc = trial(h, sk, rot);
      }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // restore: chosen work split, the saved state in the saved buffer, counters,
  // data collection and dt sequences (build_items sets n_items afresh)
  c->params = orig;
  for (auto& p : c->params) p.rows_per_wave = best_h;
  c->skew = best_skew;
  c->rotate = best_rot;
  c->desc = desc_saved;
  c->collect_nx = collect_nx_saved;
  build_items(c, best_h, c->skew, c->rotate);
  int rc2 = upload_items(c);
  for (int k = 0; k < c->n_inst; ++k) {
    const size_t n = (size_t)c->params[k].nx * c->params[k].ny;
    for (int f = 0; f < 3; ++f)
      (void)hipMemcpyAsync(c->desc[k].q[saved[k].cur][f],
                           (char*)keep[k] + (c->fp64 ? 8 : 4) * (size_t)f * n,
                           (c->fp64 ? 8 : 4) * n, hipMemcpyDeviceToDevice, c->stream);
  }
  hipError_t es = hipStreamSynchronize(c->stream);
  int rc3 = (rc || rc2 || es != hipSuccess) ? 0 : write_ctls(c, saved);
  release();
  c->unlimited = false;
  if (rc) return rc;
  if (rc2) return rc2;
  if (es != hipSuccess) return hip_fail(es, "hipStreamSynchronize (autotune)");
  if (rc3) return rc3;
  if (rows_out) *rows_out = best_h;
  return 0;
}

int fivonagi_set_skew(fivonagi_ctx* c, double skew, int rotate) {
  if (!c) return fail(FIVONAGI_ERR_ARG, "null context");
  if (!(skew < 0.0 || (skew >= 0.0 && skew <= 0.95)))
    return fail(FIVONAGI_ERR_ARG, "skew must be < 0 (automatic) or in [0, 0.95]");
  c->skew = skew < 0.0 ? -1.0f : (float)skew;
  c->rotate = rotate != 0;
  build_items(c, auto_rows(c->params, c->cus * c->per_cu), c->skew, c->rotate);
  return upload_items(c);
}

int fivonagi_enqueue(fivonagi_ctx* c, int64_t n) {
  if (!c || n < 0) return fail(FIVONAGI_ERR_ARG, "bad argument");
  if (!c->unlimited) {
    // let every instance run every attempt enqueued from now on
    std::vector<int64_t> tgt(c->n_inst, INT64_MAX);
    if (int rc = write_targets(c, tgt)) return rc;
    c->unlimited = true;
  }
  return launch_attempts(c, n);
}

int fivonagi_sync(fivonagi_ctx* c) {
  if (!c) return fail(FIVONAGI_ERR_ARG, "null context");
  std::vector<fv::Ctl> ctl;
  if (int rc = read_ctls(c, ctl)) return rc;
  for (int k = 0; k < c->n_inst; ++k) {
    if (ctl[k].err == 2) return fail(FIVONAGI_ERR_EIGEN, "imaginary eigenvalue (tmp3 < 0)");
    if (ctl[k].err == 9) return fail(FIVONAGI_ERR_DTSEQ, "prescribed dt sequence exhausted");
  }
  return 0;
}

int fivonagi_step(fivonagi_ctx* c, int64_t nsteps) {
  if (!c || nsteps < 0) return fail(FIVONAGI_ERR_ARG, "bad argument");
  std::vector<fv::Ctl> ctl;
  if (int rc = read_ctls(c, ctl)) return rc;
  std::vector<int64_t> target(c->n_inst);
  for (int k = 0; k < c->n_inst; ++k) {
    if (ctl[k].err) return fivonagi_sync(c);
    target[k] = ctl[k].n_acc + nsteps;
  }
  if (int rc = write_targets(c, target)) return rc;
  c->unlimited = false;
  int64_t remaining = nsteps;
  for (int round = 0; remaining > 0; ++round) {
    if (round > 10000) return fail(FIVONAGI_ERR_STALL, "steps keep getting rejected");
    if (int rc = launch_attempts(c, remaining)) return rc;
    if (int rc = read_ctls(c, ctl)) return rc;
    remaining = 0;
    for (int k = 0; k < c->n_inst; ++k) {
      if (ctl[k].err) return fivonagi_sync(c);
      remaining = std::max(remaining, target[k] - ctl[k].n_acc);
    }
  }
  return 0;
}

int fivonagi_step_until(fivonagi_ctx* c, double T_stop) {
  if (!c || std::isnan(T_stop)) return fail(FIVONAGI_ERR_ARG, "bad argument");
  std::vector<fv::Ctl> ctl;
  if (int rc = read_ctls(c, ctl)) return rc;
  for (int k = 0; k < c->n_inst; ++k)
    if (ctl[k].err) return fivonagi_sync(c);
  if (int rc = write_targets(c, std::vector<int64_t>(c->n_inst, INT64_MAX), T_stop)) return rc;
  c->unlimited = false;
  if (c->collect_nx > 0) {   // a new frame of the data collection
    for (int k = 0; k < c->n_inst; ++k) c->desc[k].n0frame = ctl[k].n_acc;
    if (int rc = upload_tables(c)) return rc;
  }
  // rounds of attempts sized by the slowest instance's remaining time over
  // its dt (+10%); an instance past T_stop skips the extra attempts
  for (int round = 0;; ++round) {
    int64_t need = 0;
    for (int k = 0; k < c->n_inst; ++k) {
      if (ctl[k].err) return fivonagi_sync(c);
      if (!(ctl[k].T < T_stop)) continue;
      const double steps = (T_stop - ctl[k].T) / ctl[k].dt;
      need = std::max<int64_t>(need, (int64_t)std::min(1e9, std::ceil(1.1 * steps)) + 1);
    }
    if (need == 0) break;
    if (round > 10000) return fail(FIVONAGI_ERR_STALL, "T does not reach T_stop");
    if (int rc = launch_attempts(c, std::min<int64_t>(need, 1 << 20))) return rc;
    if (int rc = read_ctls(c, ctl)) return rc;
    if (c->collect_nx > 0)
      if (int rc = collect_status(c)) return rc;
  }
  return 0;
}

int fivonagi_set_collect(fivonagi_ctx* c, int inst, const fivonagi_collect_opts* o) {
  if (int rc = check_ctx(c, inst)) return rc;
  auto& d = c->desc[inst];
  const int nx = d.nx, ny = d.ny;
  const size_t es = c->fp64 ? sizeof(double) : sizeof(float);
  const bool taylor = d.bc == fv::kBcTaylor;
  if (o) {
    if (!(o->eta > 0)) return fail(FIVONAGI_ERR_ARG, "collect: eta <= 0");
    if (!taylor && (o->jcut < 0 || o->jcut >= ny))
      return fail(FIVONAGI_ERR_ARG, "collect: jcut outside the grid");
    if (taylor && (o->cut_half < 0 || 2 * (int64_t)o->cut_half >= nx))
      return fail(FIVONAGI_ERR_ARG, "collect: the Taylor cut (2 cut_half + 1 entries) must fit a row");
    if (c->collect_bufs.size() != (size_t)c->n_inst) c->collect_bufs.resize(c->n_inst);
    auto& b = c->collect_bufs[inst];
    if (!c->d_collect_err) {
      hipError_t e;
      c->d_collect_err = (uint32_t*)c->dalloc(sizeof(uint32_t) * c->n_inst, &e);
      if (e != hipSuccess) return hip_fail(e, "hipMalloc (collect)");
      HIP_TRY(hipMemsetAsync(c->d_collect_err, 0, sizeof(uint32_t) * c->n_inst, c->stream));
    }
    if (!b.measure) {
      hipError_t e;
      b.measure = c->dalloc(es * (size_t)nx * ny, &e);
      if (e == hipSuccess) b.stage = c->dalloc(es * (size_t)nx, &e);
      if (e != hipSuccess) return hip_fail(e, "hipMalloc (collect)");
    }
    HIP_TRY(hipMemsetAsync(b.measure, 0, es * (size_t)nx * ny, c->stream));
    HIP_TRY(hipMemsetAsync(b.stage, 0, es * (size_t)nx, c->stream));
    HIP_TRY(hipMemsetAsync(c->d_collect_err + inst, 0, sizeof(uint32_t), c->stream));
    d.measure = b.measure;
    d.stage = taylor ? nullptr : b.stage;   // the Taylor cut reads the output state
    d.collect_err = c->d_collect_err + inst;
    d.jcut = o->jcut;
    d.eta = o->eta;
    for (int k = 0; k < 4; ++k) d.mic[k] = o->mic[k];
    d.cut_half = o->cut_half;
    d.tan_theta = o->tan_theta;
    c->collect_nx = std::max(c->collect_nx, nx);
    if (!taylor) c->ext_any = true;        // the staged row comes from the extended kernel
  } else {
    d.measure = nullptr;   // the buffers stay allocated for a later call
    d.stage = nullptr;
  }
  c->drop_graph();         // the captured launches change
  return upload_tables(c);
}

}  // extern "C"

// measurement grid of instance inst to a host array of type H (the
// context's precision)
template <typename H>
static int measure_impl(fivonagi_ctx* c, int inst, int rows, H* out) {
  if (int rc = check_ctx(c, inst)) return rc;
  if (!out) return fail(FIVONAGI_ERR_ARG, "null output");
  if (c->fp64 != (sizeof(H) == sizeof(double)))
    return fail(FIVONAGI_ERR_ARG, c->fp64 ? "FP64 context: use the _f64 form"
                                          : "FP32 context: use the float form");
  const auto& d = c->desc[inst];
  if (!d.measure) return fail(FIVONAGI_ERR_ARG, "data collection is off for this instance");
  if (rows < 0 || rows > d.ny) return fail(FIVONAGI_ERR_ARG, "rows outside 0..ny");
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(out, d.measure, sizeof(H) * (size_t)d.nx * rows, hipMemcpyDeviceToHost));
  return collect_status(c);
}

// The relative errors of apps/taylor-angle-xy/data_export.h:96-122 in its own
// types: float accumulators, |a - n| and the maxima in the state's type T,
// squares through pow(T, 2) (double), sqrt of the float ratio in double.
template <typename T>
static int cut_errors_impl(const T* numeric, const T* analytic, int length, float* e) {
  if (!numeric || !analytic || !e || length < 0) return fail(FIVONAGI_ERR_ARG, "cut errors");
  float l1a = 0, l1b = 0, l2a = 0, l2b = 0, lia = 0, lib = 0;
  for (int i = 0; i < length; ++i) {
    const T d = std::abs(analytic[i] - numeric[i]);
    const T a = std::abs(analytic[i]);
    l1a += d;
    l1b += a;
    l2a += (double)d * (double)d;
    l2b += (double)a * (double)a;
    lia = (float)(d > (T)lia ? d : (T)lia);
    lib = (float)(a > (T)lib ? a : (T)lib);
  }
  e[0] = l1a / l1b;
  e[1] = (float)std::sqrt((double)(l2a / l2b));
  e[2] = lia / lib;
  return 0;
}

// state of instance inst to host arrays of type H (float or double),
// converting when the context's precision differs
template <typename H>
static int collect_impl(fivonagi_ctx* c, int inst, H* q1, H* q2, H* q3, double* T, double* dt,
                        int64_t* n_acc, int64_t* n_rej) {
  if (int rc = check_ctx(c, inst)) return rc;
  std::vector<fv::Ctl> ctl;
  if (int rc = read_ctls(c, ctl)) return rc;
  const auto& t = ctl[inst];
  const auto& d = c->desc[inst];
  const size_t n = (size_t)d.nx * d.ny;
  H* dst[3] = {q1, q2, q3};
  const bool same = c->fp64 == (sizeof(H) == sizeof(double));
  std::vector<double> vd;
  std::vector<float> vf;
  for (int f = 0; f < 3; ++f) {
    if (!dst[f]) continue;
    if (same) {
      HIP_TRY(hipMemcpy(dst[f], d.q[t.cur][f], sizeof(H) * n, hipMemcpyDeviceToHost));
    } else if (c->fp64) {
      vd.resize(n);
      HIP_TRY(hipMemcpy(vd.data(), d.q[t.cur][f], sizeof(double) * n, hipMemcpyDeviceToHost));
      for (size_t o = 0; o < n; ++o) dst[f][o] = (H)vd[o];
    } else {
      vf.resize(n);
      HIP_TRY(hipMemcpy(vf.data(), d.q[t.cur][f], sizeof(float) * n, hipMemcpyDeviceToHost));
      for (size_t o = 0; o < n; ++o) dst[f][o] = (H)vf[o];
    }
  }
  if (T) *T = t.T;
  if (dt) *dt = t.dt;
  if (n_acc) *n_acc = t.n_acc;
  if (n_rej) *n_rej = t.n_rej;
  return 0;
}

extern "C" {

int fivonagi_collect(fivonagi_ctx* c, int inst, float* q1, float* q2, float* q3, double* T,
                     double* dt, int64_t* n_acc, int64_t* n_rej) {
  return collect_impl<float>(c, inst, q1, q2, q3, T, dt, n_acc, n_rej);
}

int fivonagi_collect_f64(fivonagi_ctx* c, int inst, double* q1, double* q2, double* q3,
                         double* T, double* dt, int64_t* n_acc, int64_t* n_rej) {
  return collect_impl<double>(c, inst, q1, q2, q3, T, dt, n_acc, n_rej);
}

int fivonagi_measure(fivonagi_ctx* c, int inst, float* out) {
  if (int rc = check_ctx(c, inst)) return rc;
  return measure_impl<float>(c, inst, c->desc[inst].ny, out);
}

int fivonagi_measure_f64(fivonagi_ctx* c, int inst, double* out) {
  if (int rc = check_ctx(c, inst)) return rc;
  return measure_impl<double>(c, inst, c->desc[inst].ny, out);
}

int fivonagi_measure_rows(fivonagi_ctx* c, int inst, int rows, float* out) {
  return measure_impl<float>(c, inst, rows, out);
}

int fivonagi_measure_rows_f64(fivonagi_ctx* c, int inst, int rows, double* out) {
  return measure_impl<double>(c, inst, rows, out);
}

int fivonagi_cfl_subset(int nx, int ny, uint8_t* mask) {
  if (nx < 1 || ny < 1 || !mask || (long)nx * ny > (1L << 30))
    return fail(FIVONAGI_ERR_ARG, "cfl subset: bad grid or null mask");
  const std::vector<uint8_t> m = fold_subset(nx, ny);
  std::memcpy(mask, m.data(), m.size());
  return 0;
}

int fivonagi_cut_errors(const float* numeric, const float* analytic, int length,
                        float errors[3]) {
  return cut_errors_impl<float>(numeric, analytic, length, errors);
}

int fivonagi_cut_errors_f64(const double* numeric, const double* analytic, int length,
                            float errors[3]) {
  return cut_errors_impl<double>(numeric, analytic, length, errors);
}

int fivonagi_info(fivonagi_ctx* c, int inst, fivonagi_info_t* info) {
  if (int rc = check_ctx(c, inst)) return rc;
  if (!info) return fail(FIVONAGI_ERR_ARG, "null info");
  std::vector<fv::Ctl> ctl;
  if (int rc = read_ctls(c, ctl)) return rc;
  const auto& t = ctl[inst];
  info->T = t.T;
  info->dt = t.dt;
  info->last_cfl = t.last_cfl;
  info->n_accepted = t.n_acc;
  info->n_rejected = t.n_rej;
  info->n_attempts = t.n_att;
  info->error = t.err;
  info->cur = t.cur;
  info->n_items = c->desc[inst].n_items;
  info->skew = c->skew_used;
  info->rotate = c->rotate_used ? 1 : 0;
  info->phase = (int)(c->attempts % 6);
  info->md_x = t.mdx;
  info->md_y = t.mdy;
  info->fp64 = c->fp64 ? 1 : 0;
  // mean rows per work item (heights differ when the split is skewed)
  const auto& p = c->params[inst];
  const int strips = (p.nx + fv::kStripCols - 1) / fv::kStripCols;
  const int chunks = std::max(1, c->desc[inst].n_items / strips);
  info->rows_per_wave = (int)std::lround((double)p.ny / chunks);
  return 0;
}

}  // extern "C"

namespace {

// device allocations that live for one adjoint call
// The device buffers of an adjoint call, carved from one block that the
// context keeps (fivonagi_ctx::adj_pool) and reuses on later calls: repeated
// gradients allocate nothing, and a large block freed and allocated again can
// wait for the driver to clear the freed memory.  The carving runs once with
// no base to size the block, then again to hand out the pointers; every
// buffer starts on a 256-byte boundary.
struct AdjArena {
  char* base = nullptr;
  size_t off = 0;
  template <class T>
  T* get(size_t count) {
    const size_t o = off;
    off += (sizeof(T) * std::max<size_t>(count, 1) + 255) & ~(size_t)255;
    return base ? (T*)(base + o) : nullptr;
  }
};

}  // namespace

extern "C" {

int fivonagi_step_adjoint(fivonagi_ctx* c, int64_t nsteps, const fivonagi_objective* obj,
                          const fivonagi_adjoint_opts* opts, double* J, double* grad_amp,
                          double* grad_phase) {
  if (!c || !obj || !J || nsteps <= 0) return fail(FIVONAGI_ERR_ARG, "bad argument");
  if (nsteps > (1 << 30)) return fail(FIVONAGI_ERR_ARG, "too many steps");
  if (c->md_any) return fail(FIVONAGI_ERR_ARG, "the adjoint does not support a moving window");
  if (c->fp64) return fail(FIVONAGI_ERR_ARG, "the adjoint runs on FP32 contexts");
  for (const auto& p : c->params)
    if (p.bc == FIVONAGI_BC_HIFU)
      return fail(FIVONAGI_ERR_ARG, "the adjoint does not support the HIFU boundary");
  const int ni = c->n_inst;
  for (int k = 0; k < ni; ++k) {
    const auto& p = c->params[k];
    const auto& b = obj[k];
    if (b.box_x0 < 0 || b.box_y0 < 0 || b.box_nx <= 0 || b.box_ny <= 0 ||
        b.box_x0 + b.box_nx > p.nx || b.box_y0 + b.box_ny > p.ny)
      return fail(FIVONAGI_ERR_ARG, "objective box outside the grid");
  }
  // keep every state of the forward run when asked (checkpoint_every = 1),
  // or by default when they fit in half the free device memory; else keep
  // one every K steps and recompute the segments between them
  const int req = opts ? opts->checkpoint_every : 0;
  size_t hist_bytes = 0;
  for (const auto& p : c->params)
    hist_bytes += sizeof(float) * 3 * (size_t)p.nx * p.ny * (size_t)(nsteps + 1);
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
  free_b += c->adj_pool_bytes;                 // the kept block is reused
  const bool keep_all = req == 1 || (req <= 0 && hist_bytes <= free_b / 2);
  int K = keep_all ? 64 : (req > 1 ? req : (int)std::ceil(std::sqrt((double)nsteps)));
  K = (int)std::min<int64_t>(K, nsteps);
  const int64_t ncp = (nsteps + K - 1) / K;   // segments (descriptor batches when keep_all)
  const bool plain = opts && opts->plain_forward;
  const bool check = opts && opts->check_recompute;

  std::vector<fv::Ctl> ctl;
  if (int rc = read_ctls(c, ctl)) return rc;
  for (int k = 0; k < ni; ++k)
    if (ctl[k].err) return fivonagi_sync(c);

  // ---- work items of the reverse march: strips of 56 output columns, rows
  // split so that every wave of a step is resident at once ----
  std::vector<fv::AdjItem> items;
  std::vector<size_t> first_item;
  std::vector<int32_t> wave_first;
  {
    int occ = 0, cus = 0;
    if (fv_adj_occupancy(&occ) != hipSuccess || occ <= 0) occ = 2;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) !=
            hipSuccess ||
        cus <= 0)
      cus = 256;
    fvh::build_adj_items(c->params, (long)cus * occ, opts ? opts->rows_per_wave : 0, items,
                         first_item, wave_first);
  }

  // ---- device buffers of the call ----
  std::vector<fv::AdjDesc> ad(ni);
  std::vector<float*> cps(ni);
  std::vector<fv::Ctl*> logd(ni);
  std::vector<float*> hist(ni, nullptr);
  double* jpart = nullptr;
  uint32_t* d_mis = nullptr;
  fv::AdjItem* d_items = nullptr;
  int32_t* d_wave_first = nullptr;
  fv::AdjDesc* d_ad = nullptr;
  fv::InstDesc* d_rdesc = nullptr;    // the step kernel's descriptors and control
  fv::Ctl* d_rctl = nullptr;          // blocks for the segment recompute: slot s of
  fv::SyncSlots* d_rsync = nullptr;   // every instance, and its logged control block
  auto carve = [&](AdjArena& A) {
    jpart = A.get<double>(items.size());
    d_mis = A.get<uint32_t>(1);
    for (int k = 0; k < ni; ++k) {
      const size_t n = (size_t)c->params[k].nx * c->params[k].ny;
      auto& a = ad[k];
      if (keep_all)
        hist[k] = a.seg = A.get<float>((size_t)(nsteps + 1) * 3 * n);
      else
        a.seg = A.get<float>((size_t)(K + 1) * 3 * n);
      a.lam[0] = A.get<float>(3 * n);
      a.lam[1] = A.get<float>(3 * n);
      logd[k] = A.get<fv::Ctl>(nsteps);
      a.ctl_log = logd[k];
      a.grad_ec = A.get<double>(c->params[k].ny);
      a.grad_es = A.get<double>(c->params[k].ny);
      a.jpart = jpart;
      a.mismatch = d_mis;
      cps[k] = keep_all ? nullptr : A.get<float>((size_t)ncp * 3 * n);
    }
    d_items = A.get<fv::AdjItem>(items.size());
    d_wave_first = A.get<int32_t>(wave_first.size());
    d_ad = A.get<fv::AdjDesc>(ni);
    d_rdesc = A.get<fv::InstDesc>((size_t)K * ni);
    d_rctl = A.get<fv::Ctl>((size_t)K * ni * 2);
    d_rsync = A.get<fv::SyncSlots>((size_t)3 * ni);
  };
  for (int k = 0; k < ni; ++k) {
    const auto& p = c->params[k];
    auto& a = ad[k];
    std::memset(&a, 0, sizeof(a));
    a.nx = p.nx; a.ny = p.ny; a.bc = p.bc;
    a.hbeta = 0.5f * c->desc[k].beta2m2;          // the step kernel's BETA - 1
    a.bx0 = obj[k].box_x0; a.by0 = obj[k].box_y0; a.bnx = obj[k].box_nx; a.bny = obj[k].box_ny;
    a.edge_c = c->desc[k].edge_c;
    a.edge_s = c->desc[k].edge_s;
  }
  {
    AdjArena size_only;
    carve(size_only);
    if (c->adj_pool_bytes < size_only.off) {
      HIP_TRY(hipStreamSynchronize(c->stream));
      if (c->adj_pool) (void)hipFree(c->adj_pool);
      c->adj_pool = nullptr;
      c->adj_pool_bytes = 0;
      hipError_t e = hipMalloc(&c->adj_pool, size_only.off);
      if (e != hipSuccess) {
        c->adj_pool = nullptr;
        return hip_fail(e, "hipMalloc (adjoint buffers)");
      }
      c->adj_pool_bytes = size_only.off;
    }
    AdjArena A;
    A.base = (char*)c->adj_pool;
    carve(A);
  }
  for (int k = 0; k < ni; ++k) {
    const auto& p = c->params[k];
    const size_t n = (size_t)p.nx * p.ny;
    HIP_TRY(hipMemsetAsync(ad[k].lam[0], 0, sizeof(float) * 3 * n, c->stream));
    HIP_TRY(hipMemsetAsync(ad[k].grad_ec, 0, sizeof(double) * p.ny, c->stream));
    HIP_TRY(hipMemsetAsync(ad[k].grad_es, 0, sizeof(double) * p.ny, c->stream));
    HIP_TRY(hipMemsetAsync(logd[k], 0, sizeof(fv::Ctl) * nsteps, c->stream));
  }
  HIP_TRY(hipMemsetAsync(jpart, 0, sizeof(double) * items.size(), c->stream));
  HIP_TRY(hipMemsetAsync(d_mis, 0, sizeof(uint32_t), c->stream));
  HIP_TRY(hipMemsetAsync(d_rsync, 0, sizeof(fv::SyncSlots) * 3 * ni, c->stream));
  HIP_TRY(hipMemcpyAsync(d_items, items.data(), sizeof(fv::AdjItem) * items.size(),
                         hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(d_wave_first, wave_first.data(), sizeof(int32_t) * wave_first.size(),
                         hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(d_ad, ad.data(), sizeof(fv::AdjDesc) * ni, hipMemcpyHostToDevice,
                         c->stream));

  // ---- forward: log the control block of every accepted step; keep every
  // state (the extended step kernel also stores its output into the history
  // slot of its step) or a checkpoint every K steps ----
  for (int k = 0; k < ni; ++k) {
    c->desc[k].ctl_log = logd[k];
    c->desc[k].log_base = ctl[k].n_acc;
    c->desc[k].log_len = nsteps;
    c->desc[k].hist = hist[k];
  }
  int rc = upload_tables(c);
  if (keep_all && rc == 0) {
    for (int k = 0; k < ni && rc == 0; ++k) {
      const size_t n = (size_t)c->params[k].nx * c->params[k].ny;
      for (int f = 0; f < 3 && rc == 0; ++f) {
        hipError_t e = hipMemcpyAsync(hist[k] + (size_t)f * n, c->desc[k].q[ctl[k].cur][f],
                                      sizeof(float) * n, hipMemcpyDeviceToDevice, c->stream);
        if (e != hipSuccess) rc = hip_fail(e, "state history copy");
      }
    }
    c->keep_hist = true;
    c->drop_graph();            // the captured launches become the extended kernel's
    if (rc == 0) {
      if (plain) {
        for (int64_t s = 0; s < nsteps && rc == 0; ++s) rc = fivonagi_step(c, 1);
      } else {
        rc = fivonagi_step(c, nsteps);
      }
    }
    c->keep_hist = false;
    c->drop_graph();
  }
  for (int64_t cp = 0; cp < ncp && rc == 0 && !keep_all; ++cp) {
    if ((rc = read_ctls(c, ctl))) break;
    for (int k = 0; k < ni; ++k) {
      const size_t n = (size_t)c->params[k].nx * c->params[k].ny;
      for (int f = 0; f < 3; ++f) {
        hipError_t e = hipMemcpyAsync(cps[k] + ((size_t)cp * 3 + f) * n,
                                      c->desc[k].q[ctl[k].cur][f], sizeof(float) * n,
                                      hipMemcpyDeviceToDevice, c->stream);
        if (e != hipSuccess) { rc = hip_fail(e, "checkpoint copy"); break; }
      }
    }
    if (rc) break;
    const int64_t len = std::min<int64_t>(K, nsteps - cp * K);
    if (plain) {
      // the reference's pattern: one launch per step, dt read back every step
      for (int64_t s = 0; s < len && rc == 0; ++s) rc = fivonagi_step(c, 1);
    } else {
      rc = fivonagi_step(c, len);
    }
  }
  for (int k = 0; k < ni; ++k) {
    c->desc[k].ctl_log = nullptr;
    c->desc[k].hist = nullptr;
  }
  int rc2 = upload_tables(c);
  if (rc) return rc;
  if (rc2) return rc2;
  if (int r = read_ctls(c, ctl)) return r;      // the final state's buffer (cur)
  std::vector<fv::Ctl> hlog((size_t)ni * nsteps);
  for (int k = 0; k < ni; ++k)
    HIP_TRY(hipMemcpyAsync(hlog.data() + (size_t)k * nsteps, logd[k], sizeof(fv::Ctl) * nsteps,
                           hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (size_t e = 0; e < hlog.size(); ++e)
    if (!(hlog[e].dt > 0.0)) return fail(FIVONAGI_ERR_ARG, "adjoint: step log incomplete");

  // ---- reverse, segment by segment: recompute the segment's states with
  // the step kernel (each step from its logged control block) unless every
  // state was kept, then march the adjoint back over them ----
  std::vector<fv::InstDesc> rdh((size_t)K * ni);
  std::vector<fv::Ctl> rch((size_t)K * ni * 2);
  const bool one = ni == 1;
  int par = 0;
  for (int64_t seg = ncp - 1; seg >= 0; --seg) {
    const int64_t s0 = seg * K;
    const int len = (int)std::min<int64_t>(K, nsteps - s0);
    // the previous segment's kernels read the descriptors rewritten below
    HIP_TRY(hipStreamSynchronize(c->stream));
    // the reverse kernel's slot of local step s: its own slot of the
    // segment buffer, or step s0 + s of the kept states
    const int64_t slot0 = keep_all ? s0 : 0;
    for (int k = 0; k < ni; ++k) {
      const size_t n = (size_t)c->params[k].nx * c->params[k].ny;
      float* sg = ad[k].seg + (size_t)slot0 * 3 * n;
      if (!keep_all) {
        HIP_TRY(hipMemcpyAsync(sg, cps[k] + (size_t)seg * 3 * n, sizeof(float) * 3 * n,
                               hipMemcpyDeviceToDevice, c->stream));
        // the state after the segment into slot len (rho' for the objective;
        // every field when checking the recompute against it)
        for (int f = 0; f < (check ? 3 : 1); ++f) {
          const float* src = seg + 1 < ncp ? cps[k] + ((size_t)(seg + 1) * 3 + f) * n
                                           : c->desc[k].q[ctl[k].cur][f];
          HIP_TRY(hipMemcpyAsync(sg + ((size_t)len * 3 + f) * n, src, sizeof(float) * n,
                                 hipMemcpyDeviceToDevice, c->stream));
        }
      }
      for (int s = 0; s < len; ++s) {
        fv::InstDesc d = c->desc[k];
        for (int f = 0; f < 3; ++f) {
          d.q[0][f] = sg + ((size_t)s * 3 + f) * n;
          d.q[1][f] = sg + ((size_t)(s + 1) * 3 + f) * n;
        }
        d.ctl = d_rctl + 2 * ((size_t)s * ni + k);
        d.sync = d_rsync + 3 * (size_t)k;
        d.dt_seq = nullptr;
        d.dt_seq_len = 0;
        d.ctl_log = nullptr;
        d.measure = nullptr;
        d.stage = nullptr;
        d.md_on = 0;
        rdh[(size_t)s * ni + k] = d;
        // the logged block as the attempt's own (an idle fold keeps it as is)
        fv::Ctl t = hlog[(size_t)k * nsteps + s0 + s];
        t.cur = 0;
        t.err = 0;
        t.target_acc = INT64_MAX;
        t.T_stop = INFINITY;
        rch[2 * ((size_t)s * ni + k)] = t;
        rch[2 * ((size_t)s * ni + k) + 1] = t;
      }
    }
    HIP_TRY(hipMemcpyAsync(d_rdesc, rdh.data(), sizeof(fv::InstDesc) * len * ni,
                           hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(d_rctl, rch.data(), sizeof(fv::Ctl) * 2 * len * ni,
                           hipMemcpyHostToDevice, c->stream));
    // recompute states 1 .. len-1 of the segment (phase 0: every launch
    // reads control block 0 and the never-written slot set 2, an idle fold)
    for (int s = 0; s + 1 < len && !keep_all; ++s) {
      const fv::InstDesc* di = d_rdesc + (size_t)s * ni;
      if (c->taylor_ring > 0)
        HIP_TRY(fv_launch_taylor_ghosts(di, ni, c->taylor_ring, 0, c->stream));
      HIP_TRY(fv_launch_step(di, c->d_items, c->d_rowmask, (int)c->items.size(), 0, ni,
                             one ? d_rctl + 2 * (size_t)s : nullptr, one ? d_rsync : nullptr,
                             c->params[0].bc, c->ext_any ? 1 : 0, 0, c->stream));
    }
    for (int s = len - 1; s >= 0; --s) {
      // q_n as the step read it: the Taylor ghost ring of its T (a state the
      // recompute stepped has it already; writing it again changes nothing)
      if (c->taylor_ring > 0)
        HIP_TRY(fv_launch_taylor_ghosts(d_rdesc + (size_t)s * ni, ni, c->taylor_ring, 0,
                                        c->stream));
      HIP_TRY(fv_adj_reverse(d_ad, d_items, d_wave_first, (int)wave_first.size() - 1,
                             (int)(s0 + s), (int)(slot0 + s), par, check ? 1 : 0, c->stream));
      par ^= 1;
    }
  }

  // ---- results ----
  std::vector<double> jp(items.size());
  HIP_TRY(hipMemcpyAsync(jp.data(), jpart, sizeof(double) * items.size(), hipMemcpyDeviceToHost,
                         c->stream));
  uint32_t mism = 0;
  HIP_TRY(hipMemcpyAsync(&mism, d_mis, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  size_t goff = 0;
  for (int k = 0; k < ni; ++k) {
    const auto& p = c->params[k];
    std::vector<double> gc(p.ny), gs(p.ny);
    HIP_TRY(hipMemcpyAsync(gc.data(), ad[k].grad_ec, sizeof(double) * p.ny,
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(gs.data(), ad[k].grad_es, sizeof(double) * p.ny,
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double acc = 0.0;                       // item order: the same sum every call
    for (size_t it = first_item[k]; it < first_item[k + 1]; ++it) acc += jp[it];
    J[k] = acc;
    // ampc = amp cos(phase), amps = amp sin(phase) per row
    for (int j = 0; j < p.ny; ++j) {
      const double a = c->edge_amp[k].empty() ? 0.0 : c->edge_amp[k][j];
      const double ph = c->edge_phase[k].empty() ? 0.0 : c->edge_phase[k][j];
      const double cs = std::cos(ph), sn = std::sin(ph);
      if (grad_amp) grad_amp[goff + j] = gc[j] * cs + gs[j] * sn;
      if (grad_phase) grad_phase[goff + j] = a * (gs[j] * cs - gc[j] * sn);
    }
    goff += p.ny;
  }
  if (check && mism != 0)
    return fail(FIVONAGI_ERR_CHECK, "adjoint: " + std::to_string(mism) +
                                         " recomputed cells differ from the forward's");
  // keep dJ/dq0 for fivonagi_adjoint_state (the grids of the first call are
  // reused: an instance's size does not change)
  if (c->adj_lam.size() != (size_t)ni) c->adj_lam.assign(ni, nullptr);
  for (int k = 0; k < ni; ++k) {
    const size_t n = (size_t)c->params[k].nx * c->params[k].ny;
    if (!c->adj_lam[k]) {
      hipError_t e = hipMalloc(&c->adj_lam[k], sizeof(float) * 3 * n);
      if (e != hipSuccess) {
        c->adj_lam[k] = nullptr;
        return hip_fail(e, "hipMalloc (adjoint state)");
      }
    }
    HIP_TRY(hipMemcpyAsync(c->adj_lam[k], ad[k].lam[par], sizeof(float) * 3 * n,
                           hipMemcpyDeviceToDevice, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int fivonagi_adjoint_state(fivonagi_ctx* c, int inst, float* l1, float* l2, float* l3) {
  if (int rc = check_ctx(c, inst)) return rc;
  if (c->adj_lam.empty() || !c->adj_lam[inst])
    return fail(FIVONAGI_ERR_ARG, "no adjoint has been run on this context");
  const size_t n = (size_t)c->params[inst].nx * c->params[inst].ny;
  float* dst[3] = {l1, l2, l3};
  for (int f = 0; f < 3; ++f)
    if (dst[f])
      HIP_TRY(hipMemcpy(dst[f], c->adj_lam[inst] + (size_t)f * n, sizeof(float) * n,
                        hipMemcpyDeviceToHost));
  return 0;
}

int fivonagi_step_noalloc(const fivonagi_params* p, float* const q[6], int64_t nsteps,
                          fivonagi_noalloc_state* st, void* hip_stream) {
  if (!p || !q || !st || nsteps < 0) return fail(FIVONAGI_ERR_ARG, "bad argument");
  for (int f = 0; f < 6; ++f)
    if (!q[f]) return fail(FIVONAGI_ERR_ARG, "null state pointer");
  if (int rc = check_params(*p, false)) return rc;
  if (!(st->dt > 0) || !std::isfinite(st->T)) return fail(FIVONAGI_ERR_ARG, "state: dt must be > 0");
  if (st->md_x < 0 || st->md_y < 0) return fail(FIVONAGI_ERR_ARG, "state: MDX, MDY must be >= 0");
  hipStream_t stream = (hipStream_t)hip_stream;
  void* base = nullptr;
  size_t cap = 0;
  HIP_TRY(fv_scratch_get(&base, &cap));

  // host side of the instance and its work list; `tmp` is a host-only
  // context holding what build_items reads (it owns no device memory)
  HostInst h;
  host_inst(*p, false, h);
  fivonagi_ctx tmp;
  tmp.n_inst = 1;
  tmp.params.assign(1, *p);
  tmp.desc.assign(1, h.d);
  tmp.fold_bits.assign(1, h.fold_bits);
  HIP_TRY(hipGetDevice(&tmp.device));
  device_shape(tmp.device, false, &tmp.cus, &tmp.per_cu);
  build_items(&tmp, auto_rows(tmp.params, tmp.cus * tmp.per_cu), tmp.skew, tmp.rotate);
  fv::InstDesc d = tmp.desc[0];

  // the buffer's layout, each part 256-byte aligned
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off = (off + bytes + 255) & ~(size_t)255;
    return o;
  };
  const size_t o_desc = take(sizeof(fv::InstDesc));
  const size_t o_ctl = take(2 * sizeof(fv::Ctl));
  const size_t o_sync = take(3 * sizeof(fv::SyncSlots));
  const size_t o_ec = take(sizeof(float) * h.ec.size());
  const size_t o_es = take(sizeof(float) * h.es.size());
  const size_t o_tab = take(sizeof(float) * h.tab.size());
  const size_t o_items = take(sizeof(fv::WorkItem) * tmp.items.size());
  const size_t o_mask = take(sizeof(uint64_t) * tmp.rowmask.size());
  if (off > cap)
    return fail(FIVONAGI_ERR_ARG, "fivonagi_step_noalloc: the grid's tables need " +
                                      std::to_string(off) + " bytes, more than the " +
                                      std::to_string(cap) + "-byte buffer");
  char* dev = (char*)base;
  d.q[0][0] = q[0]; d.q[0][1] = q[1]; d.q[0][2] = q[2];
  d.q[1][0] = q[3]; d.q[1][1] = q[4]; d.q[1][2] = q[5];
  d.ctl = (fv::Ctl*)(dev + o_ctl);
  d.sync = (fv::SyncSlots*)(dev + o_sync);
  d.edge_c = dev + o_ec;
  d.edge_s = dev + o_es;
  d.hifu_tab = h.tab.empty() ? nullptr : dev + o_tab;

  // control blocks of the first attempt (both blocks alike; the slot sets
  // start at zero: attempt -1 never ran)
  fv::Ctl c0;
  initial_ctl(*p, d, c0);
  c0.dt = st->dt;
  c0.T = st->T;
  c0.n_acc = st->n_accepted;
  c0.n_rej = st->n_rejected;
  c0.n_att = st->n_attempts;
  c0.mdx = st->md_x;
  c0.mdy = st->md_y;
  c0.target_acc = st->n_accepted + nsteps;
  set_ctl_derived(c0, d);

  // every table through one copy of a zero-initialised host image (the
  // reduction slots are the zeros)
  std::vector<char> img(off, 0);
  std::memcpy(img.data() + o_desc, &d, sizeof(d));
  std::memcpy(img.data() + o_ctl, &c0, sizeof(c0));
  std::memcpy(img.data() + o_ctl + sizeof(fv::Ctl), &c0, sizeof(c0));
  table_bytes(h.ec, false, img.data() + o_ec);
  table_bytes(h.es, false, img.data() + o_es);
  if (!h.tab.empty()) table_bytes(h.tab, false, img.data() + o_tab);
  std::memcpy(img.data() + o_items, tmp.items.data(), sizeof(fv::WorkItem) * tmp.items.size());
  std::memcpy(img.data() + o_mask, tmp.rowmask.data(), sizeof(uint64_t) * tmp.rowmask.size());
  HIP_TRY(hipMemcpyAsync(dev, img.data(), off, hipMemcpyHostToDevice, stream));
  HIP_TRY(hipStreamSynchronize(stream));   // img may go once the copy is done

  const fv::InstDesc* d_desc = (const fv::InstDesc*)(dev + o_desc);
  const fv::WorkItem* d_items = (const fv::WorkItem*)(dev + o_items);
  const uint64_t* d_mask = (const uint64_t*)(dev + o_mask);
  const int n_items = (int)tmp.items.size();
  const int ext = h.ext ? 1 : 0;
  int64_t a = 0;                 // attempts launched by this call (phase = a mod 6)
  fv::Ctl now = c0;
  int64_t remaining = nsteps;
  for (int round = 0; remaining > 0; ++round) {
    if (round > 10000) return fail(FIVONAGI_ERR_STALL, "steps keep getting rejected");
    for (int64_t k = 0; k < remaining; ++k, ++a) {
      const int phase = (int)(a % 6);
      if (h.taylor_ring > 0)
        HIP_TRY(fv_launch_taylor_ghosts(d_desc, 1, h.taylor_ring, phase, stream));
      HIP_TRY(fv_launch_step(d_desc, d_items, d_mask, n_items, phase, 1, d.ctl, d.sync, p->bc,
                             ext, 0, stream));
    }
    // the control block after the last attempt (fv_settle_kernel), read back
    const int phase = (int)(a % 6);
    HIP_TRY(fv_launch_settle(d_desc, 1, phase, stream));
    HIP_TRY(hipMemcpyAsync(&now, d.ctl + ((phase + 1) & 1), sizeof(fv::Ctl),
                           hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (now.err) break;
    remaining = now.target_acc - now.n_acc;
  }
  // the state ends in q[0..2]
  if (now.cur == 1)
    for (int f = 0; f < 3; ++f)
      HIP_TRY(hipMemcpyAsync(q[f], q[3 + f], sizeof(float) * (size_t)p->nx * p->ny,
                             hipMemcpyDeviceToDevice, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  st->T = now.T;
  st->dt = now.dt;
  st->n_accepted = now.n_acc;
  st->n_rejected = now.n_rej;
  st->n_attempts = now.n_att;
  st->md_x = now.mdx;
  st->md_y = now.mdy;
  st->error = now.err;
  if (now.err == 2) return fail(FIVONAGI_ERR_EIGEN, "imaginary eigenvalue (tmp3 < 0)");
  return 0;
}

void fivonagi_destroy(fivonagi_ctx* c) { delete c; }

}  // extern "C"
