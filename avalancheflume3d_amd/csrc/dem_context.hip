// dem_context.hip -- device context and C ABI (include/avflume.h) of the gfx950 DEM step.
//
// The library owns all device memory.  A context steps one or more independent instances
// (the plain method is a batch of one); instances occupy contiguous ranges of the clump
// arrays and disjoint key ranges of the sphere-cell table, so one launch of each kernel
// advances all of them.  Step sequences are captured once as hipGraphs and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "avflume.h"
#include "dem_kernels.hip"

using namespace avf;

namespace {

constexpr int kForceBlock = 64;
constexpr int kMaxNeigh = 32;
constexpr int kBlock = 256;

std::string g_create_error;

}  // namespace

struct avf_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
  avf_options opt;
  int32_t ninst = 0;
  int64_t n = 0;    // clumps over all instances
  int64_t cap = 0;  // field stride of the SoA arrays
  int64_t cell_cap = 0;
  std::vector<avf_instance_desc> desc;
  std::vector<int64_t> begin;

  InstParams *d_P = nullptr;
  double *d_st = nullptr, *d_st2 = nullptr, *d_sta = nullptr, *d_sta2 = nullptr;
  double *d_stage = nullptr, *d_f = nullptr, *d_io = nullptr, *d_io2 = nullptr;
  int32_t *d_orig = nullptr, *d_orig2 = nullptr, *d_inst = nullptr, *d_inst2 = nullptr;
  int64_t *d_inv = nullptr, *d_ncells = nullptr, *d_tmp64 = nullptr;
  uint32_t *d_counts = nullptr, *d_starts = nullptr, *d_bsum = nullptr;
  uint32_t *d_skey = nullptr, *d_srank = nullptr, *d_sid = nullptr, *d_clist = nullptr;
  int32_t *d_ccnt = nullptr;
  double4 *d_spos = nullptr;
  long long *d_acc_cell = nullptr, *d_acc_sg = nullptr;
  DevDiag *d_diag = nullptr;
  unsigned long long *d_count = nullptr;
  int64_t io_cap = 0;

  bool sgrid_valid = false;
  bool binned = false;  // sorted sphere arrays describe the current positions
  int64_t steps_since_reorder = 0;
  int64_t steps_total = 0;
  std::map<int64_t, hipGraphExec_t> graphs;
  std::vector<hipGraph_t> graph_defs;
};

namespace {

int fail(avf_ctx *ctx, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx)
    ctx->err = buf;
  else
    g_create_error = buf;
  return code;
}

#define HIPCHK(ctx, expr)                                                                  \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail((ctx), AVF_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                     \
  } while (0)

inline unsigned grid_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename T>
int dalloc(avf_ctx *ctx, T **p, int64_t count) {
  if (count <= 0) count = 1;
  hipError_t e = hipMalloc((void **)p, (size_t)count * sizeof(T));
  if (e != hipSuccess)
    return fail(ctx, AVF_ERR_NOMEM, "hipMalloc of %lld bytes failed: %s",
                (long long)(count * (int64_t)sizeof(T)), hipGetErrorString(e));
  return AVF_OK;
}

// Host mirror of InstParams for one instance.
InstParams make_params(const avf_instance_desc &d, int64_t begin) {
  InstParams p;
  std::memset(&p, 0, sizeof p);
  const avf_system &s = d.sys;
  const avf_geometry &g = d.geo;
  p.dt = s.dt;
  p.hdt = s.hdt;
  p.time = s.time;
  p.side_x = s.side.x;
  p.side_y = s.side.y;
  p.sint = s.sint;
  p.cost = s.cost;
  p.kappa_pp = s.kappa_pp;
  p.gamma_pp = s.gamma_pp;
  p.mu_pp = s.mu_pp;
  p.fr_pp = s.fr_fact_pp;
  p.kappa_pw = s.kappa_pw;
  p.gamma_pw = s.gamma_pw;
  p.mu_pw = s.mu_pw;
  p.fr_pw = s.fr_fact_pw;
  p.dy2 = s.dy2;
  p.sinfi = d.fun.sinfi;
  p.cosfi = d.fun.cosfi;
  p.tanfi = d.fun.tanfi;
  p.hop_height = g.hop_height;
  p.hop_width = g.hop_width;
  p.yfunnel = g.yfunnel;
  p.lfunnel = g.lfunnel;
  p.yhop_l = g.yhop_l;
  p.yhop_r = g.yhop_r;
  p.grav = g.grav;
  p.drag = g.drag;
  p.sphere_div = g.sphere_div;
  p.dep_sin = g.dep_sin;
  p.dep_cos = g.dep_cos;
  p.dep_x0 = g.dep_x0;
  p.dep_x1 = g.dep_x1;
  p.dep_z0 = g.dep_z0;
  p.rmax = s.rmax;
  p.gyr_max = s.gyr_max;
  p.cell_dist = g.cell_dist;
  p.regrid_ddt_steps = g.regrid_ddt_steps;
  p.begin = begin;
  p.count = s.npart;
  p.cc = 0;
  p.regrid_interval = g.regrid_interval;
  p.inter_flag = g.inter_flag;
  p.phase = d.phase;
  p.modules_fill = g.modules_fill;
  p.modules_aval = g.modules_avalanche;
  return p;
}

void reset_acc(std::vector<long long> &a, int ninst) {
  a.assign((size_t)8 * ninst, 0);
  for (int i = 0; i < ninst; i++) {
    for (int k = 0; k < 3; k++) {
      a[8 * i + k] = 0x7fffffffffffffffll;
      a[8 * i + 3 + k] = (long long)0x8000000000000000ull;
    }
    a[8 * i + 6] = (long long)0x8000000000000000ull;
    a[8 * i + 7] = 0x7fffffffffffffffll;
  }
}

// Scan of the sphere-cell counts into starts (and counts back to zero).
void launch_scan(avf_ctx *c) {
  const unsigned nb = grid_for(c->cell_cap, kScanTile);
  k_scan_a<<<nb, kScanBlock, 0, c->stream>>>(c->d_counts, c->d_bsum, c->d_ncells);
  k_scan_b<<<1, 1024, 0, c->stream>>>(c->d_bsum, c->d_ncells, c->d_starts);
  k_scan_c<<<nb, kScanBlock, 0, c->stream>>>(c->d_counts, c->d_bsum, c->d_starts, c->d_ncells);
}

// Binning of the staged spheres: predict (or stage), scan, scatter.
void launch_bin(avf_ctx *c, int mode) {
  k_predict<<<grid_for(c->n, kBlock), kBlock, 0, c->stream>>>(
      c->d_st, c->d_sta, c->d_stage, c->d_inst, c->d_orig, c->d_P, c->d_counts, c->d_skey,
      c->d_srank, c->n, c->cap, mode);
  launch_scan(c);
  k_scatter<<<std::max(1u, grid_for(4 * c->n, kBlock)), kBlock, 0, c->stream>>>(
      c->d_skey, c->d_srank, c->d_starts, c->d_stage, c->d_spos, c->d_sid, c->d_P, c->ninst,
      4 * c->n, mode);
}

template <int BS>
void launch_force_bs(avf_ctx *c, int initial) {
  k_force<kMaxNeigh, BS><<<grid_for(c->n, BS), BS, 0, c->stream>>>(
      c->d_st, c->d_sta, c->d_stage, c->d_inst, c->d_P, c->d_clist, c->d_ccnt, c->d_f, c->d_diag,
      c->n, c->cap, initial);
}

void launch_detect(avf_ctx *c) {
  k_detect<<<grid_for(4 * c->n, kBlock), kBlock, 0, c->stream>>>(
      c->d_stage, c->d_P, c->d_inst, c->d_starts, c->d_spos, c->d_sid, c->d_clist, c->d_ccnt,
      c->d_diag, 4 * c->n);
}

// k_force block shape: 64 (one wavefront) or 128 lanes (avf_options.block_clumps).
void launch_force(avf_ctx *c, int initial) {
  if (c->opt.block_clumps >= 128)
    launch_force_bs<128>(c, initial);
  else
    launch_force_bs<kForceBlock>(c, initial);
}

// One full step: avalanche_hst.cu:669-710 (or :439-467) plus the regrid and cc++.
// ev (optional, AVF_NKERNELS + 1 events) brackets every kernel for avf_step_profiled.
void launch_step(avf_ctx *c, hipEvent_t *ev = nullptr) {
  auto mark = [&](int k) {
    if (ev) (void)hipEventRecord(ev[k], c->stream);
  };
  const unsigned nb = grid_for(c->cell_cap, kScanTile);
  mark(AVF_K_PREDICT);
  k_predict<<<grid_for(c->n, kBlock), kBlock, 0, c->stream>>>(
      c->d_st, c->d_sta, c->d_stage, c->d_inst, c->d_orig, c->d_P, c->d_counts, c->d_skey,
      c->d_srank, c->n, c->cap, 1);
  mark(AVF_K_SCAN_A);
  k_scan_a<<<nb, kScanBlock, 0, c->stream>>>(c->d_counts, c->d_bsum, c->d_ncells);
  mark(AVF_K_SCAN_B);
  k_scan_b<<<1, 1024, 0, c->stream>>>(c->d_bsum, c->d_ncells, c->d_starts);
  mark(AVF_K_SCAN_C);
  k_scan_c<<<nb, kScanBlock, 0, c->stream>>>(c->d_counts, c->d_bsum, c->d_starts, c->d_ncells);
  mark(AVF_K_SCATTER);
  k_scatter<<<std::max(1u, grid_for(4 * c->n, kBlock)), kBlock, 0, c->stream>>>(
      c->d_skey, c->d_srank, c->d_starts, c->d_stage, c->d_spos, c->d_sid, c->d_P, c->ninst,
      4 * c->n, 1);
  mark(AVF_K_DETECT);
  launch_detect(c);
  mark(AVF_K_FORCE);
  launch_force(c, 0);
  mark(AVF_K_CELLREDUCE);
  k_cellreduce<<<grid_for(c->n, kBlock), kBlock, 0, c->stream>>>(c->d_st, c->d_inst, c->d_P,
                                                                  c->d_acc_cell, c->n, c->cap, 0);
  mark(AVF_K_FINALIZE);
  k_finalize<<<grid_for(c->ninst, 64), 64, 0, c->stream>>>(c->d_P, c->d_acc_cell, c->ninst, 0, 1);
  mark(AVF_NKERNELS);
}

// Sphere-cell grid from the current positions, then (optionally) the spatial reorder of the
// clump arrays by the sphere cell of each clump centre.
int rebuild_grid(avf_ctx *c, bool reorder) {
  k_sgrid_reduce<<<grid_for(c->n, kBlock), kBlock, 0, c->stream>>>(c->d_st, c->d_inst,
                                                                    c->d_acc_sg, c->n, c->cap);
  const double horizon = 2.0 * std::max(1, c->opt.reorder_interval);
  k_sgrid_finalize<<<1, 1, 0, c->stream>>>(c->d_P, c->d_acc_sg, c->ninst, c->cell_cap,
                                           c->opt.sphere_cell_factor, horizon, c->d_ncells);
  if (reorder) {
    k_reorder_key<<<grid_for(c->n, kBlock), kBlock, 0, c->stream>>>(
        c->d_st, c->d_inst, c->d_P, c->d_counts, c->d_skey, c->d_srank, c->n, c->cap);
    launch_scan(c);
    k_reorder_move<<<grid_for(c->n, kBlock), kBlock, 0, c->stream>>>(
        c->d_skey, c->d_srank, c->d_starts, c->d_st, c->d_st2, c->d_sta, c->d_sta2, c->d_orig,
        c->d_orig2, c->d_inst, c->d_inst2, c->d_P, c->d_inv, c->n, c->cap);
    const size_t nb = (size_t)c->cap * sizeof(double);
    HIPCHK(c, hipMemcpyAsync(c->d_st, c->d_st2, nb * S_NFIELDS, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_sta, c->d_sta2, nb * C_NFIELDS, hipMemcpyDeviceToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_orig, c->d_orig2, c->cap * sizeof(int32_t),
                             hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_inst, c->d_inst2, c->cap * sizeof(int32_t),
                             hipMemcpyDeviceToDevice, c->stream));
  }
  HIPCHK(c, hipGetLastError());
  c->sgrid_valid = true;
  c->steps_since_reorder = 0;
  return AVF_OK;
}

int run_steps(avf_ctx *c, int64_t k) {
  if (k <= 0) return AVF_OK;
  if (!c->opt.use_graph) {
    for (int64_t i = 0; i < k; i++) launch_step(c);
    HIPCHK(c, hipGetLastError());
    return AVF_OK;
  }
  auto it = c->graphs.find(k);
  if (it == c->graphs.end()) {
    hipGraph_t g;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int64_t i = 0; i < k; i++) launch_step(c);
    HIPCHK(c, hipStreamEndCapture(c->stream, &g));
    hipGraphExec_t ge;
    HIPCHK(c, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    c->graph_defs.push_back(g);
    it = c->graphs.emplace(k, ge).first;
  }
  HIPCHK(c, hipGraphLaunch(it->second, c->stream));
  return AVF_OK;
}

int check_inst(avf_ctx *c, int32_t inst) {
  if (!c) return AVF_ERR_ARG;
  if (inst < 0 || inst >= c->ninst) return fail(c, AVF_ERR_ARG, "instance %d out of range", inst);
  return AVF_OK;
}

int ensure_io(avf_ctx *c, int64_t count) {
  if (count <= c->io_cap) return AVF_OK;
  if (c->d_io) (void)hipFree(c->d_io);
  if (c->d_io2) (void)hipFree(c->d_io2);
  c->d_io = c->d_io2 = nullptr;
  int rc = dalloc(c, &c->d_io, count * 51);
  if (rc) return rc;
  rc = dalloc(c, &c->d_io2, count * 8);
  if (rc) return rc;
  c->io_cap = count;
  return AVF_OK;
}

}  // namespace

extern "C" {

const char *avf_last_error(const avf_ctx *ctx) {
  return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int avf_create(avf_ctx **out, const avf_instance_desc *inst, int32_t ninst,
               const avf_options *opt_in) {
  if (!out || !inst || ninst <= 0) return fail(nullptr, AVF_ERR_ARG, "bad arguments to avf_create");
  *out = nullptr;
  avf_ctx *c = new avf_ctx();
  if (opt_in)
    c->opt = *opt_in;
  else
    avf_options_default(&c->opt);
  if (c->opt.sphere_cell_factor < 1.0) c->opt.sphere_cell_factor = 1.0;
  c->opt.block_clumps = c->opt.block_clumps >= 128 ? 128 : 64;
  c->device = c->opt.device;
  c->ninst = ninst;
  c->desc.assign(inst, inst + ninst);
  int64_t total = 0;
  for (int i = 0; i < ninst; i++) {
    if (inst[i].sys.npart <= 0) {
      delete c;
      return fail(nullptr, AVF_ERR_ARG, "instance %d has npart = %lld", i,
                  (long long)inst[i].sys.npart);
    }
    c->begin.push_back(total);
    total += inst[i].sys.npart;
  }
  if (total > (int64_t)0x3fffffff) {
    delete c;
    return fail(nullptr, AVF_ERR_ARG, "too many clumps (%lld)", (long long)total);
  }
  c->n = total;
  c->cap = total;
  c->cell_cap = std::max<int64_t>(8 * 4 * total, 1 << 16);
  if (c->cell_cap > (int64_t)0xfffffff0) c->cell_cap = 0xfffffff0;

  auto bail = [&](int rc) {
    std::string e = c->err;
    avf_destroy(c);
    g_create_error = e;
    return rc;
  };
  hipError_t he = hipSetDevice(c->device);
  if (he != hipSuccess) {
    fail(c, AVF_ERR_HIP, "hipSetDevice(%d): %s", c->device, hipGetErrorString(he));
    return bail(AVF_ERR_HIP);
  }
  if (c->opt.stream) {
    c->stream = (hipStream_t)c->opt.stream;
  } else {
    he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
      fail(c, AVF_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(he));
      return bail(AVF_ERR_HIP);
    }
    c->own_stream = true;
  }
  int rc = 0;
  const int64_t N = c->cap;
  if ((rc = dalloc(c, &c->d_P, ninst)) || (rc = dalloc(c, &c->d_st, N * S_NFIELDS)) ||
      (rc = dalloc(c, &c->d_st2, N * S_NFIELDS)) || (rc = dalloc(c, &c->d_sta, N * C_NFIELDS)) ||
      (rc = dalloc(c, &c->d_sta2, N * C_NFIELDS)) ||
      (rc = dalloc(c, &c->d_stage, N * kStageDoubles)) || (rc = dalloc(c, &c->d_f, N * F_NFIELDS)) ||
      (rc = dalloc(c, &c->d_orig, N)) || (rc = dalloc(c, &c->d_orig2, N)) ||
      (rc = dalloc(c, &c->d_inst, N)) || (rc = dalloc(c, &c->d_inst2, N)) ||
      (rc = dalloc(c, &c->d_inv, N)) || (rc = dalloc(c, &c->d_ncells, 1)) ||
      (rc = dalloc(c, &c->d_tmp64, N)) || (rc = dalloc(c, &c->d_counts, c->cell_cap)) ||
      (rc = dalloc(c, &c->d_starts, c->cell_cap + 1)) ||
      (rc = dalloc(c, &c->d_bsum, grid_for(c->cell_cap, kScanTile) + 1)) ||
      (rc = dalloc(c, &c->d_skey, 4 * N)) || (rc = dalloc(c, &c->d_srank, 4 * N)) ||
      (rc = dalloc(c, &c->d_sid, 4 * N)) || (rc = dalloc(c, &c->d_spos, 4 * N)) ||
      (rc = dalloc(c, &c->d_clist, 4 * N * kSphList)) || (rc = dalloc(c, &c->d_ccnt, 4 * N)) ||
      (rc = dalloc(c, &c->d_acc_cell, 8 * ninst)) || (rc = dalloc(c, &c->d_acc_sg, 8 * ninst)) ||
      (rc = dalloc(c, &c->d_diag, 1)) || (rc = dalloc(c, &c->d_count, 1)))
    return bail(rc);

  // Parameters, identity storage order, zeroed tables.
  std::vector<InstParams> hp;
  std::vector<int32_t> orig(N), instv(N);
  std::vector<int64_t> inv(N);
  for (int i = 0; i < ninst; i++) {
    InstParams p = make_params(inst[i], c->begin[i]);
    // reach of a sphere centre from the clump centre, plus the sphere radius
    p.sg_rsph = inst[i].sys.rmax;
    p.sg_reach = inst[i].sys.rmax * (std::sqrt(3.0) / inst[i].geo.sphere_div) + inst[i].sys.rmax;
    hp.push_back(p);
    for (int64_t k = 0; k < inst[i].sys.npart; k++) {
      orig[c->begin[i] + k] = (int32_t)k;
      instv[c->begin[i] + k] = i;
      inv[c->begin[i] + k] = c->begin[i] + k;
    }
  }
  std::vector<long long> acc;
  reset_acc(acc, ninst);
#define UP(dst, src, bytes) \
  if ((he = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)) != hipSuccess) { \
    fail(c, AVF_ERR_HIP, "upload: %s", hipGetErrorString(he));                 \
    return bail(AVF_ERR_HIP);                                                   \
  }
  UP(c->d_P, hp.data(), hp.size() * sizeof(InstParams));
  UP(c->d_orig, orig.data(), N * sizeof(int32_t));
  UP(c->d_inst, instv.data(), N * sizeof(int32_t));
  UP(c->d_inv, inv.data(), N * sizeof(int64_t));
  UP(c->d_acc_cell, acc.data(), acc.size() * sizeof(long long));
  UP(c->d_acc_sg, acc.data(), acc.size() * sizeof(long long));
#undef UP
  if ((he = hipMemset(c->d_counts, 0, c->cell_cap * sizeof(uint32_t))) != hipSuccess ||
      (he = hipMemset(c->d_st, 0, N * S_NFIELDS * sizeof(double))) != hipSuccess ||
      (he = hipMemset(c->d_sta, 0, N * C_NFIELDS * sizeof(double))) != hipSuccess ||
      (he = hipMemset(c->d_f, 0, N * F_NFIELDS * sizeof(double))) != hipSuccess ||
      (he = hipMemset(c->d_diag, 0, sizeof(DevDiag))) != hipSuccess ||
      (he = hipMemset(c->d_ncells, 0, sizeof(int64_t))) != hipSuccess) {
    fail(c, AVF_ERR_HIP, "hipMemset: %s", hipGetErrorString(he));
    return bail(AVF_ERR_HIP);
  }
  *out = c;
  return AVF_OK;
}

int avf_destroy(avf_ctx *c) {
  if (!c) return AVF_OK;
  for (auto &kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto g : c->graph_defs) (void)hipGraphDestroy(g);
  void *bufs[] = {c->d_P,      c->d_st,    c->d_st2,   c->d_sta,   c->d_sta2,  c->d_stage,
                  c->d_f,      c->d_io,    c->d_io2,   c->d_orig,  c->d_orig2, c->d_inst,
                  c->d_inst2,  c->d_inv,   c->d_ncells, c->d_tmp64, c->d_counts, c->d_starts,
                  c->d_bsum,   c->d_skey,  c->d_srank, c->d_sid,   c->d_spos,  c->d_acc_cell,
                  c->d_acc_sg, c->d_diag,  c->d_count, c->d_clist, c->d_ccnt};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return AVF_OK;
}

int avf_upload(avf_ctx *c, int32_t inst, const avf_particle *pp, const avf_double3 *aa_i,
               const avf_double3 *anga_i) {
  int rc = check_inst(c, inst);
  if (rc) return rc;
  const int64_t cnt = c->desc[inst].sys.npart;
  // a NULL argument keeps the device's current values
  std::vector<avf_particle> cur;
  std::vector<avf_double3> ca, cg;
  if (!pp || !aa_i || !anga_i) {
    cur.resize((size_t)cnt);
    ca.resize((size_t)cnt);
    cg.resize((size_t)cnt);
    if ((rc = avf_download(c, inst, cur.data(), ca.data(), cg.data()))) return rc;
    if (!pp) pp = cur.data();
    if (!aa_i) aa_i = ca.data();
    if (!anga_i) anga_i = cg.data();
  }
  if ((rc = ensure_io(c, cnt))) return rc;
  std::vector<double> extra((size_t)cnt * 8, 0.0);
  const double div = c->desc[inst].geo.sphere_div;
  for (int64_t k = 0; k < cnt; k++) {
    double *x = &extra[8 * k];
    x[0] = aa_i[k].x;
    x[1] = aa_i[k].y;
    x[2] = aa_i[k].z;
    x[3] = anga_i[k].x;
    x[4] = anga_i[k].y;
    x[5] = anga_i[k].z;
    // pre-freeze radius: the record's, or for a frozen clump (radius 0, init_Verlet.cu:31)
    // the one its stale sphere centre implies
    double r = pp[k].radius;
    if (r == 0.0) {
      const double dx = pp[k].rr_a.x - pp[k].rr.x, dy = pp[k].rr_a.y - pp[k].rr.y;
      const double dz = pp[k].rr_a.z - pp[k].rr.z;
      r = std::sqrt(dx * dx + dy * dy + dz * dz) / std::sqrt(3.0) * div;
    }
    x[6] = r;
  }
  HIPCHK(c, hipMemcpyAsync(c->d_io, pp, cnt * sizeof(avf_particle), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_io2, extra.data(), extra.size() * sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
  k_import_aos<<<grid_for(cnt, kBlock), kBlock, 0, c->stream>>>(
      c->d_st, c->d_sta, c->d_inv, c->begin[inst], c->d_io, c->d_io2, cnt, c->cap);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->sgrid_valid = false;
  c->binned = false;
  return AVF_OK;
}

int avf_download(avf_ctx *c, int32_t inst, avf_particle *pp, avf_double3 *aa_i,
                 avf_double3 *anga_i) {
  int rc = check_inst(c, inst);
  if (rc) return rc;
  const int64_t cnt = c->desc[inst].sys.npart;
  if ((rc = ensure_io(c, cnt))) return rc;
  const bool want_a = aa_i || anga_i;
  k_export_aos<<<grid_for(c->n, kBlock), kBlock, 0, c->stream>>>(
      c->d_st, c->d_sta, c->d_orig, c->d_inst, c->d_P, inst, c->d_io, want_a ? c->d_io2 : nullptr,
      nullptr, c->d_f, c->n, c->cap);
  HIPCHK(c, hipGetLastError());
  if (pp)
    HIPCHK(c, hipMemcpyAsync(pp, c->d_io, cnt * sizeof(avf_particle), hipMemcpyDeviceToHost,
                             c->stream));
  std::vector<double> a;
  if (want_a) {
    a.resize((size_t)cnt * 6);
    HIPCHK(c, hipMemcpyAsync(a.data(), c->d_io2, a.size() * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int64_t k = 0; k < (want_a ? cnt : 0); k++) {
    if (aa_i) aa_i[k] = {a[6 * k], a[6 * k + 1], a[6 * k + 2]};
    if (anga_i) anga_i[k] = {a[6 * k + 3], a[6 * k + 4], a[6 * k + 5]};
  }
  return AVF_OK;
}

int avf_get_forces(avf_ctx *c, int32_t inst, avf_double3 *force, avf_double3 *torque) {
  int rc = check_inst(c, inst);
  if (rc) return rc;
  const int64_t cnt = c->desc[inst].sys.npart;
  if ((rc = ensure_io(c, cnt))) return rc;
  k_export_aos<<<grid_for(c->n, kBlock), kBlock, 0, c->stream>>>(
      c->d_st, c->d_sta, c->d_orig, c->d_inst, c->d_P, inst, c->d_io, nullptr, c->d_io2, c->d_f,
      c->n, c->cap);
  HIPCHK(c, hipGetLastError());
  std::vector<double> f((size_t)cnt * 6);
  HIPCHK(c, hipMemcpyAsync(f.data(), c->d_io2, f.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int64_t k = 0; k < cnt; k++) {
    if (force) force[k] = {f[6 * k], f[6 * k + 1], f[6 * k + 2]};
    if (torque) torque[k] = {f[6 * k + 3], f[6 * k + 4], f[6 * k + 5]};
  }
  return AVF_OK;
}

int avf_set_phase(avf_ctx *c, int32_t inst, int32_t phase, double time, double side_x, int64_t cc) {
  int rc = check_inst(c, inst);
  if (rc) return rc;
  InstParams p;
  HIPCHK(c, hipMemcpyAsync(&p, c->d_P + inst, sizeof p, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  p.phase = phase;
  p.time = time;
  p.side_x = side_x;
  p.cc = cc;
  HIPCHK(c, hipMemcpyAsync(c->d_P + inst, &p, sizeof p, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->desc[inst].phase = phase;
  return AVF_OK;
}

int avf_get_time(avf_ctx *c, int32_t inst, double *time, int64_t *cc) {
  int rc = check_inst(c, inst);
  if (rc) return rc;
  InstParams p;
  HIPCHK(c, hipMemcpyAsync(&p, c->d_P + inst, sizeof p, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (time) *time = p.time;
  if (cc) *cc = p.cc;
  return AVF_OK;
}

int avf_regrid(avf_ctx *c) {
  if (!c) return AVF_ERR_ARG;
  k_cellreduce<<<grid_for(c->n, kBlock), kBlock, 0, c->stream>>>(c->d_st, c->d_inst, c->d_P,
                                                                  c->d_acc_cell, c->n, c->cap, 1);
  k_finalize<<<grid_for(c->ninst, 64), 64, 0, c->stream>>>(c->d_P, c->d_acc_cell, c->ninst, 1, 0);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AVF_OK;
}

int avf_get_cell(avf_ctx *c, int32_t inst, avf_cell *cell) {
  int rc = check_inst(c, inst);
  if (rc) return rc;
  InstParams p;
  HIPCHK(c, hipMemcpyAsync(&p, c->d_P + inst, sizeof p, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  cell->ntags = c->desc[inst].ntags;
  cell->ncell = {p.ncell[0], p.ncell[1], p.ncell[2]};
  cell->ncell3 = p.ncell[0] * p.ncell[1] * p.ncell[2];
  cell->ntot = cell->ncell3 * cell->ntags;
  cell->cell_side = {p.cell_side[0], p.cell_side[1], p.cell_side[2]};
  cell->cs_inv = {p.cs_inv[0], p.cs_inv[1], p.cs_inv[2]};
  cell->rr_max = {p.rr_max[0], p.rr_max[1], p.rr_max[2]};
  cell->rr_min = {p.rr_min[0], p.rr_min[1], p.rr_min[2]};
  cell->length = {p.length[0], p.length[1], p.length[2]};
  return AVF_OK;
}

int avf_initial_forces(avf_ctx *c) {
  if (!c) return AVF_ERR_ARG;
  int rc;
  if (!c->sgrid_valid && (rc = rebuild_grid(c, c->opt.reorder_interval > 0))) return rc;
  launch_bin(c, 0);
  launch_detect(c);
  launch_force(c, 1);
  c->binned = true;
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AVF_OK;
}

int avf_step(avf_ctx *c, int64_t nsteps) {
  if (!c || nsteps < 0) return AVF_ERR_ARG;
  int rc;
  const int64_t chunk_max = c->opt.reorder_interval > 0 ? c->opt.reorder_interval : 1000;
  int64_t left = nsteps;
  while (left > 0) {
    if (!c->sgrid_valid ||
        (c->opt.reorder_interval > 0 && c->steps_since_reorder >= c->opt.reorder_interval)) {
      if ((rc = rebuild_grid(c, c->opt.reorder_interval > 0))) return rc;
    }
    int64_t k = std::min(left, chunk_max - (c->steps_since_reorder % chunk_max));
    if ((rc = run_steps(c, k))) return rc;
    c->binned = true;
    c->steps_since_reorder += k;
    c->steps_total += k;
    left -= k;
  }
  return AVF_OK;
}

int avf_step_profiled(avf_ctx *c, int64_t nsteps, double *kernel_ms, double *total_ms) {
  if (!c || nsteps < 0) return AVF_ERR_ARG;
  // Two batches of events: batch b+1 is queued before batch b is read, so reading the
  // timings never leaves the GPU idle.
  constexpr int kBatch = 128, kEv = AVF_NKERNELS + 1;
  std::vector<hipEvent_t> ev((size_t)2 * kBatch * kEv);
  for (auto &e : ev) HIPCHK(c, hipEventCreate(&e));
  hipEvent_t first, last;
  HIPCHK(c, hipEventCreate(&first));
  HIPCHK(c, hipEventCreate(&last));
  double sum[AVF_NKERNELS] = {0};
  int rc = AVF_OK;
  int64_t queued = 0;
  int64_t batch_len[2] = {0, 0};
  int cur = 0;
  bool started = false;
  auto read_batch = [&](int b) -> int {
    if (batch_len[b] == 0) return AVF_OK;
    hipEvent_t *base = &ev[(size_t)b * kBatch * kEv];
    HIPCHK(c, hipEventSynchronize(base[(batch_len[b] - 1) * kEv + AVF_NKERNELS]));
    for (int64_t i = 0; i < batch_len[b]; i++)
      for (int j = 0; j < AVF_NKERNELS; j++) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, base[i * kEv + j], base[i * kEv + j + 1]));
        sum[j] += ms;
      }
    batch_len[b] = 0;
    return AVF_OK;
  };
  while (queued < nsteps && rc == AVF_OK) {
    const int64_t k = std::min<int64_t>(kBatch, nsteps - queued);
    hipEvent_t *base = &ev[(size_t)cur * kBatch * kEv];
    for (int64_t i = 0; i < k; i++) {
      if (!c->sgrid_valid ||
          (c->opt.reorder_interval > 0 && c->steps_since_reorder >= c->opt.reorder_interval)) {
        if ((rc = rebuild_grid(c, c->opt.reorder_interval > 0))) break;
      }
      if (!started) {
        HIPCHK(c, hipEventRecord(first, c->stream));
        started = true;
      }
      launch_step(c, &base[i * kEv]);
      c->steps_since_reorder++;
      c->steps_total++;
      batch_len[cur]++;
    }
    HIPCHK(c, hipGetLastError());
    queued += k;
    cur ^= 1;
    if ((rc = read_batch(cur))) break;  // the previous batch, while this one runs
  }
  if (rc == AVF_OK) rc = read_batch(cur ^ 1);
  HIPCHK(c, hipEventRecord(last, c->stream));
  HIPCHK(c, hipEventSynchronize(last));
  float tot = 0.f;
  if (started) HIPCHK(c, hipEventElapsedTime(&tot, first, last));
  for (auto &e : ev) (void)hipEventDestroy(e);
  (void)hipEventDestroy(first);
  (void)hipEventDestroy(last);
  c->binned = true;
  for (int j = 0; j < AVF_NKERNELS; j++)
    if (kernel_ms) kernel_ms[j] = nsteps ? sum[j] / (double)nsteps : 0.0;
  if (total_ms) *total_ms = tot;
  return rc;
}

int avf_sync(avf_ctx *c) {
  if (!c) return AVF_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AVF_OK;
}

int avf_get_clump_cells(avf_ctx *c, int32_t inst, int64_t *cell_index) {
  int rc = check_inst(c, inst);
  if (rc) return rc;
  const int64_t cnt = c->desc[inst].sys.npart;
  k_export_cells<<<grid_for(c->n, kBlock), kBlock, 0, c->stream>>>(c->d_st, c->d_orig, c->d_inst,
                                                                    c->d_P, inst, c->d_tmp64, c->n,
                                                                    c->cap);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(cell_index, c->d_tmp64, cnt * sizeof(int64_t), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AVF_OK;
}

// cells.cu:79-102 on the current state: the first clump (ascending index) that makes
// cell_locate exit, and why.
int avf_check_cells(avf_ctx *c, int32_t inst, int64_t *bad) {
  int rc = check_inst(c, inst);
  if (rc) return rc;
  const int64_t cnt = c->desc[inst].sys.npart, ntags = c->desc[inst].ntags;
  std::vector<int64_t> cells((size_t)cnt);
  if ((rc = avf_get_clump_cells(c, inst, cells.data()))) return rc;
  std::map<int64_t, int64_t> occ;
  for (int64_t k = 0; k < cnt; k++) {
    if (cells[k] == -1) continue;
    if (cells[k] == -2) {
      if (bad) *bad = k;
      return AVF_ERR_CELL_NEG;
    }
    int64_t &o = occ[cells[k]];
    if (o == ntags - 1) {
      if (bad) *bad = k;
      return AVF_ERR_CELL_FULL;
    }
    o++;
  }
  if (bad) *bad = -1;
  return AVF_OK;
}

int avf_get_contacts(avf_ctx *c, int32_t inst, avf_contact *out, int64_t capn, int64_t *count) {
  int rc = check_inst(c, inst);
  if (rc) return rc;
  if (!c->binned)
    return fail(c, AVF_ERR_ARG, "no force stage since the last upload (call avf_initial_forces)");
  int64_t room = std::max<int64_t>(capn, 0);
  int64_t *d_out = nullptr;
  if ((rc = dalloc(c, &d_out, 4 * std::max<int64_t>(room, 1)))) return rc;
  HIPCHK(c, hipMemsetAsync(c->d_count, 0, sizeof(unsigned long long), c->stream));
  k_export_contacts<<<grid_for(c->n, kBlock), kBlock, 0, c->stream>>>(
      c->d_st, c->d_sta, c->d_stage, c->d_orig, c->d_inst, c->d_P, inst, c->d_starts, c->d_spos,
      c->d_sid, d_out, c->d_count, room, c->n, c->cap);
  hipError_t he = hipGetLastError();
  unsigned long long total = 0;
  std::vector<int64_t> raw;
  if (he == hipSuccess)
    he = hipMemcpyAsync(&total, c->d_count, sizeof total, hipMemcpyDeviceToHost, c->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
  const int64_t got = std::min<int64_t>((int64_t)total, room);
  raw.resize((size_t)4 * std::max<int64_t>(got, 1));
  if (he == hipSuccess && got > 0)
    he = hipMemcpy(raw.data(), d_out, (size_t)got * 4 * sizeof(int64_t), hipMemcpyDeviceToHost);
  (void)hipFree(d_out);
  if (he != hipSuccess) return fail(c, AVF_ERR_HIP, "contact export: %s", hipGetErrorString(he));
  std::vector<avf_contact> v((size_t)got);
  for (int64_t k = 0; k < got; k++) {
    v[k].l = raw[4 * k];
    v[k].m = raw[4 * k + 1];
    v[k].sl = (int32_t)(raw[4 * k + 2] & 0xff);
    v[k].sm = (int32_t)(raw[4 * k + 2] >> 8);
    v[k].repulsive = (int32_t)raw[4 * k + 3];
    v[k].reserved = 0;
  }
  std::sort(v.begin(), v.end(), [](const avf_contact &a, const avf_contact &b) {
    if (a.l != b.l) return a.l < b.l;
    if (a.sm != b.sm) return a.sm < b.sm;
    if (a.sl != b.sl) return a.sl < b.sl;
    return a.m < b.m;
  });
  if (out && got > 0) std::memcpy(out, v.data(), (size_t)got * sizeof(avf_contact));
  if (count) *count = (int64_t)total;
  return AVF_OK;
}

int avf_observe(avf_ctx *c, int32_t inst, avf_observables *obs, const avf_profile_spec *spec,
                double *profile) {
  int rc = check_inst(c, inst);
  if (rc) return rc;
  const int64_t cnt = c->desc[inst].sys.npart;
  int nbins = 0;
  double x0 = 0.0, bw = 1.0, sigma = 1.0, width = 1.0;
  if (spec && profile) {
    if (spec->nbins <= 0 || spec->nbins > kMaxBins || !(spec->x1 > spec->x0) ||
        !(spec->width > 0.0))
      return fail(c, AVF_ERR_ARG, "bad deposit profile spec (nbins 1..%d, x1 > x0, width > 0)",
                  kMaxBins);
    nbins = spec->nbins;
    x0 = spec->x0;
    bw = (spec->x1 - spec->x0) / nbins;
    sigma = spec->sigma > 0.0 ? spec->sigma : bw;
    width = spec->width;
  }
  const int64_t nb = (cnt + kObsBlock - 1) / kObsBlock;
  const int64_t w = kObsScalars + nbins;
  double *d_part = nullptr, *d_out = nullptr;
  if ((rc = dalloc(c, &d_part, nb * w)) || (rc = dalloc(c, &d_out, w))) {
    if (d_part) (void)hipFree(d_part);
    return rc;
  }
  k_observe<<<(unsigned)nb, kObsBlock, 0, c->stream>>>(c->d_st, c->d_sta, c->d_inv, c->d_P, inst,
                                                      cnt, c->cap, nbins, x0, bw, sigma, width,
                                                      d_part);
  k_observe_final<<<1, 1, 0, c->stream>>>(d_part, nb, nbins, d_out);
  std::vector<double> h((size_t)w);
  hipError_t he = hipGetLastError();
  if (he == hipSuccess)
    he = hipMemcpyAsync(h.data(), d_out, w * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
  (void)hipFree(d_part);
  (void)hipFree(d_out);
  if (he != hipSuccess) return fail(c, AVF_ERR_HIP, "observe: %s", hipGetErrorString(he));
  if (obs) {
    obs->vmean = h[0] / (double)cnt;  // avalanche_hst.cu:573-574 (vmean /= nn)
    obs->wmean = h[1] / (double)cnt;
    obs->xcm = h[2];
    obs->zcm = h[3];
    // xcm /= tot_mass (avalanche_hst.cu:827-828); tot_mass from the instance description or,
    // when it was not given, the sum of the clump masses
    double tot = c->desc[inst].sys.tot_mass;
    if (!(tot > 0.0)) tot = h[8];
    obs->xcm /= tot;
    obs->zcm /= tot;
    obs->runout_mean = h[6] > 0.0 ? h[4] / h[6] : 0.0;
    obs->runout_front = h[5];
    obs->mass_live = h[6];
    obs->live = (int64_t)h[7];
  }
  for (int b = 0; b < nbins; b++) profile[b] = h[kObsScalars + b];
  return AVF_OK;
}

int avf_get_diag(avf_ctx *c, avf_diag *d) {
  if (!c || !d) return AVF_ERR_ARG;
  DevDiag h;
  HIPCHK(c, hipMemcpyAsync(&h, c->d_diag, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_diag, 0, sizeof h, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  d->cell_below = (int64_t)h.cell_below;
  d->cell_beyond = (int64_t)h.cell_beyond;
  d->neigh_overflow = (int64_t)h.neigh_overflow;
  d->max_neighbours = (int64_t)h.max_neighbours;
  d->max_cell_spheres = (int64_t)h.max_cell;
  d->steps = c->steps_total;
  return AVF_OK;
}

}  // extern "C"
