// dem_context.hip -- device context and C ABI (include/avflume.h) of the gfx950 DEM step.
//
// The library owns all device memory.  A context steps one or more independent instances
// (the plain method is a batch of one); instances occupy contiguous ranges of the clump
// arrays and disjoint key ranges of the sphere-cell table, so one launch of each kernel
// advances all of them.  Step sequences are captured once as hipGraphs and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "avflume.h"
#include "dem_kernels.hip"
#include "dem_window.hip"

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
  double *d_stage2 = nullptr;  // second stage buffer of the fused list steps
  bool lean = false;           // avf_options.lean_state resolved at create
  bool closed = false;         // avf_options.closed_forms resolved at create
  bool deep = false;           // k_force gathers candidates ahead (small contexts)
  std::vector<hipEvent_t> prof_ev;  // timing events of the profiled / timed steps
  int32_t *d_orig = nullptr, *d_orig2 = nullptr, *d_inst = nullptr, *d_inst2 = nullptr;
  int64_t *d_inv = nullptr, *d_ncells = nullptr, *d_nbricks = nullptr, *d_tmp64 = nullptr;
  uint32_t *d_counts = nullptr, *d_starts = nullptr, *d_bsum = nullptr;
  uint32_t *d_bsum2 = nullptr;  // per-step binning: share sums added up by the count pass (x2)
  uint32_t *d_skey = nullptr, *d_srank = nullptr, *d_sid = nullptr, *d_clist = nullptr;
  int32_t *d_ccnt = nullptr;
  double4 *d_spos = nullptr;
  // per-step binning by clump (avf_options.step_binning 0): one key, rank and sorted
  // record per clump, the clump grid's cell and brick counts, the records' largest
  // coordinate; the lists go to the candidate-list arrays below
  bool clump_bins = false;
  uint32_t *d_ckey = nullptr, *d_crank = nullptr;
  uint2 *d_cik = nullptr;  // storage index and key of each clump in cell order
  float4 *d_frec = nullptr;  // float records of the clump sweep (centre, radius), storage order
  float4 *d_sca = nullptr;      // their centres and radii in cell order
  int64_t *d_cgrid = nullptr;  // [0] clump cells over all instances
  // candidate lists (verlet_skin > 0): per clump up to kVMax neighbour clumps with their
  // candidate sphere pairings (entry-major), displacement tracking, chunk checkpoint
  int32_t *d_vl_m = nullptr, *d_vl_cnt = nullptr;
  uint16_t *d_vl_mask = nullptr;
  double *d_dsp = nullptr;
  uint8_t *d_wreach = nullptr;
  double *d_ck_st = nullptr, *d_ck_rad = nullptr;
  InstParams *d_ck_P = nullptr;
  DevDiag *d_ck_diag = nullptr;
  unsigned long long *d_umax = nullptr;
  // pinned host words for the small read-backs of the step loops (the chunk's displacement
  // bound, the diagnostic counters, the instances' parameters at a run's start)
  struct HostMail {
    unsigned long long umax;
    DevDiag diag;
    double obs[kObsScalars + kMaxBins];  // avf_observe_noalloc's sums
  } *h_mail = nullptr;
  double *d_scratch = nullptr;  // this device's address of scratch (avf_observe_noalloc)
  InstParams *h_P = nullptr;
  unsigned long long seen_below = 0, seen_overflow = 0;  // counters at the last check
  unsigned long long seen_broken = 0;
  bool list_valid = false;
  int64_t vchunk = 8;             // next chunk of candidate-list steps
  int64_t since_build = 0;        // steps since the last list build
  long long *d_acc_cell = nullptr, *d_acc_sg = nullptr;
  DevDiag *d_diag = nullptr;
  unsigned long long *d_count = nullptr;
  int64_t io_cap = 0;

  bool sgrid_valid = false;
  bool binned = false;   // sorted sphere arrays describe the current positions
  bool stepped = false;  // st holds the state after steps (the exports may bin it)
  int64_t steps_since_reorder = 0;
  int64_t buf_set = 0;  // which of the two state buffer sets is current (swapped by reorders)
  int64_t steps_total = 0;
  int64_t list_builds = 0, steps_undone = 0;  // avf_diag
  std::map<int64_t, hipGraphExec_t> graphs;
  std::vector<hipGraph_t> graph_defs;
  struct WinBuf *win = nullptr;  // buffers of the differentiable window (on first use)
  int32_t prof_kernel = -1;      // avf_step_timed: the one kernel bracketed by events
  int32_t prof_every = 1;        // avf_step_timed: events on every prof_every-th step
  int64_t prof_seq = 0;          // steps launched under the profiler
};

// Block partial sums and totals of avf_observe_noalloc: a fixed device array of
// kObsScratchDoubles doubles (8 MiB), one copy on each GPU the library runs on, held from
// the code object's load until the process ends.  Every call on a GPU shares that GPU's copy
// (calls that overlap on one GPU race); it keeps the last call's contents, and nothing in it
// needs to start at zero (k_observe writes every partial that k_observe_final reads).  Host
// code uses the address hipGetSymbolAddress gives for the current device, never the name.
constexpr int64_t kObsScratchDoubles = int64_t{1} << 20;
static __device__ __align__(256) double scratch[kObsScratchDoubles];

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

// scratch's address on each device, looked up once per device by the first context created
// on it (that device current).
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kScratchDevices] = {};

int scratch_address(avf_ctx *c) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (c->device < 0 || c->device >= kScratchDevices)
    return fail(c, AVF_ERR_ARG, "device %d out of range", c->device);
  if (!g_scratch[c->device]) {
    void *p = nullptr;
    const hipError_t he = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (he != hipSuccess || !p)
      return fail(c, AVF_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(he));
    g_scratch[c->device] = static_cast<double *>(p);
  }
  c->d_scratch = g_scratch[c->device];
  return AVF_OK;
}

// One k_copy_set launch: {source, destination, 4-byte words} per range (unused ranges empty).
struct CopyRange {
  const void *src;
  void *dst;
  int64_t words;
};
void copy_set(avf_ctx *c, const std::array<CopyRange, kCopySets> &r) {
  CopySet cs{};
  int64_t most = 0;
  for (int k = 0; k < kCopySets; k++) {
    cs.src[k] = static_cast<const uint32_t *>(r[k].src);
    cs.dst[k] = static_cast<uint32_t *>(r[k].dst);
    cs.words[k] = r[k].words;
    most = std::max(most, r[k].words / 4 + 1);
  }
  k_copy_set<<<std::min<unsigned>(kCopyGrid, grid_for(most, kCopyBlock)), kCopyBlock, 0,
               c->stream>>>(cs);
}

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

// Workgroups of the two scan launches (each takes whole tiles of kScanTile cells).
unsigned scan_grid(const avf_ctx *c) {
  return std::max(1u, std::min((unsigned)kScanGrid, grid_for(c->cell_cap, kScanTile)));
}

// Scan of the sphere-cell counts into starts (counts back to zero).  finalize_prev /
// advance: the parameter updates between two force stages of per-step binning
// (step_params: the previous step's regrid and cc++, the next step's time advance).
void launch_scan(avf_ctx *c, uint32_t *starts, int finalize_prev = 0, int advance = 0) {
  const unsigned G = scan_grid(c);
  k_count_sum<<<G, kScanBlock, 0, c->stream>>>(c->d_counts, c->d_bsum, c->d_ncells);
  k_count_scan<<<G, kScanBlock, 0, c->stream>>>(c->d_counts, c->d_bsum, starts,
                                                c->d_ncells, c->d_P, c->d_acc_cell, c->ninst,
                                                finalize_prev, advance, 1, nullptr);
}

// Spheres into cell order at their cell's start (from the scan) + their rank (from the
// count).  ccnt: the per-sphere counts of the sweep that follows (those of clumps not binned
// are zeroed here); NULL for a binning that no sweep reads.
void launch_scatter(avf_ctx *c, const uint32_t *skey, const uint32_t *srank,
                    const uint32_t *starts, const double *stage, double4 *spos, uint32_t *sid,
                    int32_t *ccnt) {
  k_scatter<<<std::max(1u, grid_for(4 * c->n, kScatterBlock)), kScatterBlock, 0, c->stream>>>(
      skey, srank, starts, stage, spos, sid, c->d_inst, c->d_P, ccnt, c->d_diag, 4 * c->n);
}

// Binning of the staged spheres: predict (mode 1, advancing the time) or stage (mode 0),
// scan, scatter.
void launch_bin(avf_ctx *c, int mode) {
  k_predict<<<grid_for(c->n, kPredictBlock), kPredictBlock, 0, c->stream>>>(
      c->d_st, c->d_sta, c->d_stage, c->d_inst, c->d_orig, c->d_P, c->d_counts, c->d_skey,
      c->d_srank, c->n, c->cap, mode, ListTrack{nullptr, nullptr}, 0, nullptr);
  launch_scan(c, c->d_starts, 0, mode == 1);
  launch_scatter(c, c->d_skey, c->d_srank, c->d_starts, c->d_stage, c->d_spos, c->d_sid,
                 c->d_ccnt);
}

ListTrack list_track(avf_ctx *c) { return ListTrack{c->d_dsp, c->d_wreach}; }

template <int BS, bool VERLET, bool NEXT, bool LEAN = false, bool CLOSED = false,
          bool DEEP = false, bool CELLS = false>
void launch_force_bs(avf_ctx *c, int initial, bool ft, int time_ahead, const double *stage_in,
                     double *stage_out, int si, uint32_t *bsum = nullptr) {
  k_force<kMaxNeigh, BS, VERLET, NEXT, LEAN, CLOSED, DEEP, CELLS>
      <<<grid_for(c->n, BS), BS, 0, c->stream>>>(
      c->d_st, c->d_sta, stage_in, c->d_inst, c->d_P, c->d_clist, c->d_ccnt, c->d_vl_m,
      c->d_vl_mask, c->d_vl_cnt, ft ? c->d_f : nullptr, c->d_diag, c->d_acc_cell, stage_out,
      c->d_orig, (VERLET && !CELLS) ? list_track(c) : ListTrack{nullptr, nullptr}, c->n, c->cap,
      initial, time_ahead, si, c->d_counts, (VERLET && CELLS) ? c->d_ckey : c->d_skey,
      (VERLET && CELLS) ? c->d_crank : c->d_srank, (VERLET && CELLS) ? c->d_frec : nullptr,
      bsum, (int32_t)scan_grid(c), (VERLET && CELLS) ? c->d_cgrid : c->d_ncells);
}

// Touching spheres of the binned spheres (skin 0) or candidate spheres within sep + skin.
void launch_detect(avf_ctx *c, double skin = 0.0) {
  if (skin > 0.0)
    k_sweep<kSphListV><<<kSwGrid, kSwBlock, 0, c->stream>>>(
        c->d_P, c->ninst, c->d_nbricks, c->d_starts, c->d_spos, c->d_sid, c->d_clist,
        c->d_ccnt, c->d_diag, skin);
  else
    k_sweep<kSphList><<<kSwGrid, kSwBlock, 0, c->stream>>>(
        c->d_P, c->ninst, c->d_nbricks, c->d_starts, c->d_spos, c->d_sid, c->d_clist,
        c->d_ccnt, c->d_diag, 0.0);
}

// k_force block shape: 64 (one wavefront) or 128 lanes (avf_options.block_clumps).  ft:
// keep the step's forces and torques for avf_get_forces (the last step of a sequence).
// verlet: touching clumps from the candidate lists; time_ahead: see k_force.  stage_in: the
// step's stage records (default d_stage); stage_out (candidate lists only): also predict the
// next step into that buffer.  si: step index within a chunk of candidate-list steps.
void launch_force(avf_ctx *c, int initial, bool ft = true, bool verlet = false,
                  int time_ahead = 0, const double *stage_in = nullptr,
                  double *stage_out = nullptr, int si = 0) {
  if (!stage_in) stage_in = c->d_stage;
  const bool next = verlet && stage_out;
  if (c->opt.block_clumps >= 128) {
    if (next && c->lean)
      launch_force_bs<128, true, true, true>(c, initial, ft, time_ahead, stage_in, stage_out,
                                             si);
    else if (next)
      launch_force_bs<128, true, true>(c, initial, ft, time_ahead, stage_in, stage_out, si);
    else if (verlet)
      launch_force_bs<128, true, false>(c, initial, ft, time_ahead, stage_in, nullptr, si);
    else
      launch_force_bs<128, false, false>(c, initial, ft, time_ahead, stage_in, nullptr, si);
  } else {
    if (next && c->lean && c->closed)
      launch_force_bs<kForceBlock, true, true, true, true>(c, initial, ft, time_ahead, stage_in,
                                                           stage_out, si);
    else if (next && c->lean && c->deep)
      launch_force_bs<kForceBlock, true, true, true, false, true>(c, initial, ft, time_ahead,
                                                                  stage_in, stage_out, si);
    else if (next && c->lean)
      launch_force_bs<kForceBlock, true, true, true>(c, initial, ft, time_ahead, stage_in,
                                                     stage_out, si);
    else if (next)
      launch_force_bs<kForceBlock, true, true>(c, initial, ft, time_ahead, stage_in, stage_out,
                                               si);
    else if (verlet && c->closed)
      launch_force_bs<kForceBlock, true, false, false, true>(c, initial, ft, time_ahead,
                                                             stage_in, nullptr, si);
    else if (verlet && c->deep)
      launch_force_bs<kForceBlock, true, false, false, false, true>(c, initial, ft, time_ahead,
                                                                    stage_in, nullptr, si);
    else if (verlet)
      launch_force_bs<kForceBlock, true, false>(c, initial, ft, time_ahead, stage_in, nullptr,
                                                si);
    else
      launch_force_bs<kForceBlock, false, false>(c, initial, ft, time_ahead, stage_in, nullptr,
                                                 si);
  }
}

// The force stage of per-step binning: own state from the stage record (stage_in), touching
// spheres from the sweep; with stage_out, the next step's predict into that buffer with its
// sphere-cell counts and ranks, otherwise the state back to st.
void launch_cforce(avf_ctx *c, bool ft, const double *stage_in, double *stage_out,
                   uint32_t *bsum_next = nullptr) {
  if (c->clump_bins) {  // touching clumps from the clump sweep's lists; with a next step, its
                        // counts, ranks and share sums too
    if (stage_out && c->closed)
      launch_force_bs<kForceBlock, true, true, true, true, false, true>(c, 0, ft, 0, stage_in,
                                                                        stage_out, 0, bsum_next);
    else if (stage_out)
      launch_force_bs<kForceBlock, true, true, true, false, false, true>(c, 0, ft, 0, stage_in,
                                                                         stage_out, 0, bsum_next);
    else if (c->closed)
      launch_force_bs<kForceBlock, true, false, true, true, false, true>(c, 0, ft, 0, stage_in,
                                                                         nullptr, 0);
    else
      launch_force_bs<kForceBlock, true, false, true, false, false, true>(c, 0, ft, 0, stage_in,
                                                                          nullptr, 0);
    return;
  }
  if (stage_out && c->closed)
    launch_force_bs<kForceBlock, false, true, true, true, false, true>(c, 0, ft, 0, stage_in,
                                                                       stage_out, 0);
  else if (stage_out)
    launch_force_bs<kForceBlock, false, true, true, false, false, true>(c, 0, ft, 0, stage_in,
                                                                        stage_out, 0);
  else if (c->closed)
    launch_force_bs<kForceBlock, false, false, true, true, false, true>(c, 0, ft, 0, stage_in,
                                                                        nullptr, 0);
  else
    launch_force_bs<kForceBlock, false, false, true, false, false, true>(c, 0, ft, 0, stage_in,
                                                                         nullptr, 0);
}

// Step i of a chunk of S steps with the sphere cells rebuilt every step
// (avalanche_hst.cu:669-710 or :439-467, plus the regrid and cc++).  The chunk's first step
// predicts from st (k_predict, keys, counts and ranks); every step then scans the counts
// (applying the previous step's regrid and cc++ and advancing the time), places the spheres
// in cell order, sweeps the cells for touching spheres and runs the force stage, which also
// predicts the next step into the other stage buffer with its keys, and k_count_ranks counts
// those keys in this step's sorted order; the chunk's last force stage writes the state
// back to st instead and one k_finalize applies its regrid and cc++.  ev (optional,
// AVF_NKERNELS + 1 events) brackets the kernels in the slots of avf_step_profiled (SCAN_A:
// count sums of a chunk's first step, SCAN_B: scan, SCAN_C: empty, DETECT: sweep, COUNT: the
// next step's counts with the scan's share sums).
void launch_cstep(avf_ctx *c, int64_t i, int64_t S, hipEvent_t *ev = nullptr, bool ft = true) {
  auto mark = [&](int k) {
    if (ev && (c->prof_kernel < 0 || k == c->prof_kernel || k == c->prof_kernel + 1))
      (void)hipEventRecord(ev[k], c->stream);
  };
  double *buf[2] = {c->d_stage, c->d_stage2};
  const double *cur = buf[i & 1];
  const unsigned G = scan_grid(c);
  // the keys, ranks and cell count of the binning: one per clump in the clump grid, or
  // four per clump in the sphere grid
  const bool cb = c->clump_bins;
  uint32_t *key = cb ? c->d_ckey : c->d_skey, *rank = cb ? c->d_crank : c->d_srank;
  int64_t *ncells = cb ? c->d_cgrid : c->d_ncells;
  mark(AVF_K_PREDICT);
  if (i == 0)
    k_predict<<<grid_for(c->n, kPredictBlock), kPredictBlock, 0, c->stream>>>(
        c->d_st, c->d_sta, c->d_stage, c->d_inst, c->d_orig, c->d_P, c->d_counts, key, rank,
        c->n, c->cap, 3, ListTrack{nullptr, nullptr}, cb ? 1 : 0, cb ? c->d_frec : nullptr);
  mark(AVF_K_SCAN_A);
  // the share sums of the scan: k_count_sum on a chunk's first step, after that added up by
  // the previous step's count pass into the step's half of d_bsum2 (the scan clears the
  // other half for the next count)
  uint32_t *bs = i == 0 ? c->d_bsum : c->d_bsum2 + (size_t)(i & 1) * kScanGrid * kBsumStride;
  uint32_t *bs_next = c->d_bsum2 + (size_t)((i + 1) & 1) * kScanGrid * kBsumStride;
  if (i == 0) k_count_sum<<<G, kScanBlock, 0, c->stream>>>(c->d_counts, c->d_bsum, ncells);
  mark(AVF_K_SCAN_B);
  k_count_scan<<<G, kScanBlock, 0, c->stream>>>(c->d_counts, bs, c->d_starts, ncells, c->d_P,
                                                c->d_acc_cell, c->ninst, i > 0, 1,
                                                i == 0 ? 1 : kBsumStride, bs_next);
  mark(AVF_K_SCAN_C);
  mark(AVF_K_SCATTER);
  if (cb)
    k_cscatter<<<std::max(1u, grid_for(c->n, 256)), 256, 0, c->stream>>>(
        c->d_ckey, c->d_crank, c->d_starts, c->d_frec, c->d_sca, c->d_cik,
        c->d_vl_cnt, c->d_diag, c->n);
  else
    launch_scatter(c, c->d_skey, c->d_srank, c->d_starts, cur, c->d_spos, c->d_sid, c->d_ccnt);
  mark(AVF_K_DETECT);
  if (cb)
    k_csweep<<<std::max(1u, grid_for(c->n, kCsBlock)), kCsBlock, 0, c->stream>>>(
        c->d_P, c->ninst, c->d_cgrid, c->d_starts, c->d_sca, cur, c->d_cik,
        c->d_vl_m, c->d_vl_mask, c->d_vl_cnt, c->d_diag, c->cap, c->n);
  else
    launch_detect(c);
  mark(AVF_K_FORCE);
  launch_cforce(c, ft, cur, i + 1 < S ? buf[(i + 1) & 1] : nullptr, bs_next);
  mark(AVF_K_CELLREDUCE);  // the regrid reduction runs in k_force
  mark(AVF_K_FINALIZE);
  if (i + 1 == S)
    k_finalize<<<grid_for(c->ninst, 64), 64, 0, c->stream>>>(c->d_P, c->d_acc_cell, c->ninst, 0,
                                                             1, 0, 1);
  mark(AVF_K_COUNT);
  if (i + 1 < S && !cb)  // the next step's counts and ranks, in this step's sorted order (by
                         // clump, the force kernel counts them)
    k_count_ranks<<<std::max(1u, grid_for(4 * c->n, 256)), 256, 0, c->stream>>>(
        key, c->d_sid, c->d_starts, ncells, c->d_counts, rank, bs_next, (int32_t)G, 4 * c->n);
  mark(AVF_NKERNELS);
}

// One step on candidate lists.  rebuild: bin the predicted spheres and rebuild the lists
// (k_predict keys, scans, scatter advancing P.time, k_detect with the skin, k_vmerge), then
// the force stage.  Otherwise step i of a chunk of S list steps: k_predict without keys
// (tracking displacements) for the first step only; k_force runs the force stage at the
// step's time + dt and, except on the last step, predicts the next step into the other
// stage buffer; after the last step one k_finalize advances P by the chunk's S steps (the
// caller ends a chunk on any step whose regrid is due).  The kernel slots of
// avf_step_profiled are kept (the binning slots are empty on list steps, the predict slot
// after the first step of a chunk, the finalize slot before its last).
void launch_vstep(avf_ctx *c, bool rebuild, hipEvent_t *ev = nullptr, bool ft = true,
                  int64_t i = 0, int64_t S = 1, unsigned marks = ~0u) {
  auto mark = [&](int k) {  // marks: which of the boundaries record their event
    if (ev && ((marks >> k) & 1u) &&
        (c->prof_kernel < 0 || k == c->prof_kernel || k == c->prof_kernel + 1))
      (void)hipEventRecord(ev[k], c->stream);
  };
  const unsigned G = scan_grid(c);
  mark(AVF_K_PREDICT);
  if (rebuild || i == 0)
    k_predict<<<grid_for(c->n, kPredictBlock), kPredictBlock, 0, c->stream>>>(
        c->d_st, c->d_sta, c->d_stage, c->d_inst, c->d_orig, c->d_P, c->d_counts, c->d_skey,
        c->d_srank, c->n, c->cap, rebuild ? 1 : 2, list_track(c), 0, nullptr);
  mark(AVF_K_SCAN_A);
  if (rebuild) k_count_sum<<<G, kScanBlock, 0, c->stream>>>(c->d_counts, c->d_bsum, c->d_ncells);
  mark(AVF_K_SCAN_B);
  if (rebuild)  // also advances the time (between the predict and the force stage)
    k_count_scan<<<G, kScanBlock, 0, c->stream>>>(c->d_counts, c->d_bsum, c->d_starts,
                                                  c->d_ncells, c->d_P,
                                                  c->d_acc_cell, c->ninst, 0, 1, 1, nullptr);
  mark(AVF_K_SCAN_C);
  mark(AVF_K_SCATTER);
  if (rebuild)
    launch_scatter(c, c->d_skey, c->d_srank, c->d_starts, c->d_stage, c->d_spos, c->d_sid,
                   c->d_ccnt);
  mark(AVF_K_DETECT);
  if (rebuild) {
    launch_detect(c, c->opt.verlet_skin);
    k_vmerge<kForceBlock><<<grid_for(c->n, kForceBlock), kForceBlock, 0, c->stream>>>(
        c->d_stage, c->d_P, c->d_inst, c->d_orig, c->d_clist, c->d_ccnt, c->d_vl_m, c->d_vl_mask,
        c->d_vl_cnt, list_track(c), c->d_diag, c->n, c->cap, c->opt.verlet_skin);
  }
  mark(AVF_K_FORCE);
  if (rebuild) {
    launch_force(c, 0, ft, true, 0);
  } else {
    double *buf[2] = {c->d_stage, c->d_stage2};
    launch_force(c, 0, ft, true, 1, buf[i & 1], i + 1 < S ? buf[(i + 1) & 1] : nullptr, (int)i);
  }
  mark(AVF_K_CELLREDUCE);  // the regrid reduction runs in k_force
  mark(AVF_K_FINALIZE);
  if (rebuild)
    k_finalize<<<grid_for(c->ninst, 64), 64, 0, c->stream>>>(c->d_P, c->d_acc_cell, c->ninst, 0,
                                                             1, 0, 1);
  else if (i + 1 == S)
    k_finalize<<<grid_for(c->ninst, 64), 64, 0, c->stream>>>(c->d_P, c->d_acc_cell, c->ninst, 0,
                                                             1, 1, (int)S);
  mark(AVF_K_COUNT);  // per-step binning only
  mark(AVF_NKERNELS);
}

// Sphere-cell grid from the current positions, then (optionally) the spatial reorder of the
// clump arrays by the sphere cell of each clump centre.
int rebuild_grid(avf_ctx *c, bool reorder) {
  k_sgrid_reduce<<<grid_for(c->n, kBlock * kSgItems), kBlock, 0, c->stream>>>(
      c->d_st, c->d_inst, c->d_acc_sg, c->n, c->cap);
  const double horizon = 2.0 * std::max(1, c->opt.reorder_interval);
  k_sgrid_finalize<<<1, 1, 0, c->stream>>>(c->d_P, c->d_acc_sg, c->ninst, c->cell_cap,
                                           c->opt.sphere_cell_factor, horizon, c->d_ncells,
                                           c->opt.verlet_skin, c->d_nbricks,
                                           c->clump_bins ? c->d_cgrid : nullptr);
  if (reorder) {
    k_reorder_key<<<grid_for(c->n, kBlock), kBlock, 0, c->stream>>>(
        c->d_st, c->d_inst, c->d_P, c->d_counts, c->d_skey, c->d_srank, c->n, c->cap);
    launch_scan(c, c->d_starts);
    k_reorder_move<<<grid_for(c->n, kBlock), kBlock, 0, c->stream>>>(
        c->d_skey, c->d_srank, c->d_starts, c->d_st, c->d_st2, c->d_sta, c->d_sta2, c->d_orig,
        c->d_orig2, c->d_inst, c->d_inst2, c->d_P, c->d_inv, c->n, c->cap);
    // the moved arrays become the current ones (captured graphs are kept per buffer set)
    std::swap(c->d_st, c->d_st2);
    std::swap(c->d_sta, c->d_sta2);
    std::swap(c->d_orig, c->d_orig2);
    std::swap(c->d_inst, c->d_inst2);
    c->buf_set ^= 1;
  }
  HIPCHK(c, hipGetLastError());
  c->sgrid_valid = true;
  c->list_valid = false;  // storage order (and the sphere grid) changed
  c->binned = false;      // the sorted sphere arrays describe the old grid and order
  c->steps_since_reorder = 0;
  return AVF_OK;
}

// Per-kernel event sums of avf_step_profiled.
struct StepProf {
  double sum[AVF_NKERNELS] = {0};
  int64_t count[AVF_NKERNELS] = {0};  // timed launches per slot (per-step binning)
  int64_t launched = 0;
};

// At least n timing events, created once per context and reused: timing-only, so without
// the system-scope cache writeback / invalidate at each record.
int prof_events(avf_ctx *c, size_t n, hipEvent_t **out) {
  while (c->prof_ev.size() < n) {
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess)
      return fail(c, AVF_ERR_HIP, "hipEventCreateWithFlags failed");
    c->prof_ev.push_back(e);
  }
  *out = c->prof_ev.data();
  return AVF_OK;
}

// S candidate-list (or one rebuild) steps launched eagerly with events between kernels.
int vsteps_profiled(avf_ctx *c, int64_t S, bool rebuild, StepProf *prof) {
  constexpr int kEv = AVF_NKERNELS + 1;
  hipEvent_t *ev = nullptr;
  int rc = prof_events(c, (size_t)S * kEv, &ev);
  const int64_t E = c->prof_every;
  if (rc == AVF_OK && !rebuild && c->prof_kernel == AVF_K_FORCE && E > 1) {
    // k_force of a chunk's list steps: runs of E consecutive launches (back to back, nothing
    // else between them) inside one pair of events, one run starting every 2E steps; never
    // the chunk's last step (another force variant).  The events' own cost is then shared
    // by E launches.
    std::vector<int64_t> runs;  // first step of each run
    int64_t run_end = -1;
    for (int64_t i = 0; i < S; i++) {
      unsigned marks = 0u;
      if (run_end < 0 && c->prof_seq % (2 * E) == 0 && i + E <= S - 1) {
        runs.push_back(i);
        run_end = i + E - 1;
        marks |= 1u << AVF_K_FORCE;
      }
      if (i == run_end) marks |= 1u << (AVF_K_FORCE + 1);
      c->prof_seq++;
      launch_vstep(c, rebuild, marks ? &ev[(runs.size() - 1) * kEv] : nullptr, i + 1 == S, i,
                   S, marks);
      if (i == run_end) run_end = -1;
    }
    hipError_t he = hipSuccess;
    if (!runs.empty()) he = hipEventSynchronize(ev[(runs.size() - 1) * kEv + AVF_K_FORCE + 1]);
    for (size_t r = 0; r < runs.size() && he == hipSuccess; r++) {
      float ms = 0.f;
      he = hipEventElapsedTime(&ms, ev[r * kEv + AVF_K_FORCE], ev[r * kEv + AVF_K_FORCE + 1]);
      prof->sum[AVF_K_FORCE] += ms;
      prof->launched += E;
    }
    if (he != hipSuccess) rc = fail(c, AVF_ERR_HIP, "event timing: %s", hipGetErrorString(he));
    return rc;
  }
  if (rc == AVF_OK) {
    std::vector<int64_t> timed;  // steps of this chunk carrying events
    for (int64_t i = 0; i < S; i++) {
      // (runs of k_force list steps time only that variant: rebuild steps stay out)
      const bool rec = (c->prof_every <= 1 || c->prof_seq % c->prof_every == 0) &&
                       !(rebuild && c->prof_kernel == AVF_K_FORCE && c->prof_every > 1);
      c->prof_seq++;
      if (rec) timed.push_back(i);
      launch_vstep(c, rebuild, rec ? &ev[(size_t)i * kEv] : nullptr, i + 1 == S, i, S);
    }
    const int jlast = c->prof_kernel >= 0 ? c->prof_kernel + 1 : AVF_NKERNELS;
    hipError_t he = hipSuccess;
    if (!timed.empty()) he = hipEventSynchronize(ev[(size_t)timed.back() * kEv + jlast]);
    for (size_t t = 0; t < timed.size() && he == hipSuccess; t++)
      for (int j = 0; j < AVF_NKERNELS && he == hipSuccess; j++) {
        if (c->prof_kernel >= 0 && j != c->prof_kernel) continue;
        const size_t i = (size_t)timed[t];
        float ms = 0.f;
        he = hipEventElapsedTime(&ms, ev[i * kEv + j], ev[i * kEv + j + 1]);
        prof->sum[j] += ms;
      }
    if (he != hipSuccess) rc = fail(c, AVF_ERR_HIP, "event timing: %s", hipGetErrorString(he));
    prof->launched += (int64_t)timed.size();
  }
  return rc;
}

// Steps on candidate lists (avf_options.verlet_skin > 0).  A rebuild step bins the spheres
// and lists every sphere pair within sep + skin; chunks of list steps follow, each from a
// checkpoint of the state.  After a chunk the largest displacement bound of any sphere since
// the build decides: beyond 0.49 skin a pair outside the list could have touched, so the
// chunk is undone from its checkpoint and redone after a rebuild; otherwise the next chunk
// length is set from the bound's growth.  The results are those of per-step binning.
// Status of the steps just taken, from the device's counters (compared with their values
// at the previous check; avf_get_diag restarts them): a clump outside its instance's
// clump-cell grid, where the reference's cell_locate exits (cells.cu:79-88) or writes past
// its table, and a neighbour list that overflowed (contacts dropped).
int step_status(avf_ctx *c) {
  HIPCHK(c, hipMemcpyAsync(&c->h_mail->diag, c->d_diag, sizeof(DevDiag), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const unsigned long long below = c->h_mail->diag.cell_below;
  const unsigned long long over = c->h_mail->diag.neigh_overflow;
  const unsigned long long nb = below - c->seen_below, no = over - c->seen_overflow;
  const unsigned long long broken = c->h_mail->diag.bin_broken - c->seen_broken;
  c->seen_below = below;
  c->seen_overflow = over;
  c->seen_broken = c->h_mail->diag.bin_broken;
  if (broken)
    return fail(c, AVF_ERR_HIP, "%llu binned entries beyond the binned count (binning broken)",
                broken);
  if (nb)
    return fail(c, AVF_ERR_CELL_NEG,
                "%llu clump visit(s) outside the clump-cell grid (the reference exits)", nb);
  if (no)
    return fail(c, AVF_ERR_NEIGH_OVERFLOW, "%llu neighbour list overflow(s): contacts dropped",
                no);
  return AVF_OK;
}

int run_verlet(avf_ctx *c, int64_t nsteps, StepProf *prof) {
  // chunk length cap: one chunk per reorder interval where the displacement bound allows
  // (each chunk costs a checkpoint, its first predict, its last step and a host round trip)
  constexpr int64_t kMaxChunk = 512, kGraphSteps = 128;
  const double skin = c->opt.verlet_skin;
  const size_t SF = (size_t)c->cap * S_NFIELDS * sizeof(double);
  int rc;
  int64_t left = nsteps;
  // clump-cell regrid schedule: a chunk's k_finalize applies its steps at its end, so every
  // chunk ends on a step whose regrid is due (cc % regrid_interval == 0), if it has one
  HIPCHK(c, hipMemcpyAsync(c->h_P, c->d_P, c->ninst * sizeof(InstParams),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<InstParams> hp(c->h_P, c->h_P + c->ninst);
  auto to_regrid = [&hp]() {  // steps up to and including the next due one
    int64_t lim = std::numeric_limits<int64_t>::max();
    for (const InstParams &q : hp)
      if (q.regrid_interval > 0) {
        const int64_t iv = q.regrid_interval, r = ((q.cc % iv) + iv) % iv;
        lim = std::min<int64_t>(lim, (iv - r) % iv + 1);
      }
    return lim;
  };
  auto advance = [&hp](int64_t k) {
    for (InstParams &q : hp) q.cc += k;
  };
  while (left > 0) {
    if (!c->sgrid_valid ||
        (c->opt.reorder_interval > 0 && c->steps_since_reorder >= c->opt.reorder_interval)) {
      if ((rc = rebuild_grid(c, c->opt.reorder_interval > 0))) return rc;
    }
    if (!c->list_valid) {
      if (prof) {
        if ((rc = vsteps_profiled(c, 1, true, prof))) return rc;
      } else {
        launch_vstep(c, true);
      }
      c->list_valid = true;
      c->binned = true;
      c->since_build = 0;
      c->list_builds++;
      c->steps_since_reorder++;
      c->steps_total++;
      advance(1);
      left--;
      continue;
    }
    int64_t S = std::min(left, std::max<int64_t>(1, c->vchunk));
    if (c->opt.reorder_interval > 0)
      S = std::min<int64_t>(S, c->opt.reorder_interval - c->steps_since_reorder);
    S = std::min(S, to_regrid());
    // checkpoint: state, radii (clumps may freeze), parameters (time, cc, grids), counters
    copy_set(c, {{{c->d_st, c->d_ck_st, (int64_t)(SF / 4)},
                  {c->d_sta + (size_t)C_RAD * c->cap, c->d_ck_rad, c->cap * 2},
                  {c->d_P, c->d_ck_P, c->ninst * (int64_t)(sizeof(InstParams) / 4)},
                  {c->d_diag, c->d_ck_diag, (int64_t)(sizeof(DevDiag) / 4)}}});
    if (prof) {
      if ((rc = vsteps_profiled(c, S, false, prof))) return rc;
    } else if (c->opt.use_graph) {
      // the chunk as a sequence of graphs of at most kGraphSteps steps (long graphs replay
      // slower than a few short ones).  Only the piece holding the chunk's last step depends
      // on the chunk length (its force variant, k_finalize over S steps): the others are
      // cached by their first step alone.  Plain chunks use positive keys.
      for (int64_t off = 0; off < S; off += kGraphSteps) {
        const int64_t len = std::min(kGraphSteps, S - off);
        const int64_t key = (off + len < S ? -(int64_t{1} << 40) - off
                                           : -(S * (kMaxChunk + 1) + off + 1)) ^
                            (c->buf_set << 56);
        auto it = c->graphs.find(key);
        if (it == c->graphs.end()) {
          hipGraph_t g;
          HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
          for (int64_t i = off; i < off + len; i++)
            launch_vstep(c, false, nullptr, i + 1 == S, i, S);
          HIPCHK(c, hipStreamEndCapture(c->stream, &g));
          hipGraphExec_t ge;
          HIPCHK(c, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
          c->graph_defs.push_back(g);
          it = c->graphs.emplace(key, ge).first;
        }
        HIPCHK(c, hipGraphLaunch(it->second, c->stream));
      }
    } else {
      for (int64_t i = 0; i < S; i++) launch_vstep(c, false, nullptr, i + 1 == S, i, S);
    }
    // largest displacement bound since the build
    HIPCHK(c, hipMemsetAsync(c->d_umax, 0, sizeof(unsigned long long), c->stream));
    k_win_max<<<std::min(kWinMaxGrid, grid_for(c->n, kWinMaxBlock)), kWinMaxBlock, 0,
                c->stream>>>(c->d_dsp, c->n, c->d_umax);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(&c->h_mail->umax, c->d_umax, sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double maxd;
    std::memcpy(&maxd, &c->h_mail->umax, sizeof maxd);
    if (maxd > 0.49 * skin) {
      copy_set(c, {{{c->d_ck_st, c->d_st, (int64_t)(SF / 4)},
                    {c->d_ck_rad, c->d_sta + (size_t)C_RAD * c->cap, c->cap * 2},
                    {c->d_ck_P, c->d_P, c->ninst * (int64_t)(sizeof(InstParams) / 4)},
                    {c->d_ck_diag, c->d_diag, (int64_t)(sizeof(DevDiag) / 4)}}});
      c->list_valid = false;
      c->vchunk = std::max<int64_t>(1, S / 2);
      c->steps_undone += S;
      continue;
    }
    left -= S;
    advance(S);
    c->steps_since_reorder += S;
    c->steps_total += S;
    c->since_build += S;
    c->binned = false;
    c->stepped = true;
    const double rate = maxd / (double)c->since_build, room = 0.49 * skin - maxd;
    int64_t next = rate > 0.0 ? (int64_t)(0.8 * room / rate) : kMaxChunk;
    if (next < 1) {
      c->list_valid = false;
      next = 1;
    }
    c->vchunk = std::min(next, kMaxChunk);
    if (c->vchunk > 32) c->vchunk -= c->vchunk % 32;  // few distinct lengths: few graphs
  }
  return AVF_OK;
}

// S steps of per-step binning launched eagerly with events between the kernels of the
// steps that are timed (every prof_every-th; for avf_step_timed only the brackets of
// prof_kernel).
int csteps_profiled(avf_ctx *c, int64_t S, StepProf *prof) {
  constexpr int kEv = AVF_NKERNELS + 1;
  hipEvent_t *ev = nullptr;
  int rc = prof_events(c, (size_t)S * kEv, &ev);
  if (rc) return rc;
  std::vector<int64_t> timed;
  for (int64_t i = 0; i < S; i++) {
    const bool rec = c->prof_every <= 1 || c->prof_seq % c->prof_every == 0;
    c->prof_seq++;
    if (rec) timed.push_back(i);
    launch_cstep(c, i, S, rec ? &ev[(size_t)i * kEv] : nullptr, i + 1 == S);
  }
  const int jlast = c->prof_kernel >= 0 ? c->prof_kernel + 1 : AVF_NKERNELS;
  hipError_t he = hipSuccess;
  if (!timed.empty()) he = hipEventSynchronize(ev[(size_t)timed.back() * kEv + jlast]);
  for (size_t t = 0; t < timed.size() && he == hipSuccess; t++)
    for (int j = 0; j < AVF_NKERNELS && he == hipSuccess; j++) {
      if (c->prof_kernel >= 0 && j != c->prof_kernel) continue;
      // slots of kernels this step does not launch stay out of the sums
      if (j == AVF_K_SCAN_C || j == AVF_K_CELLREDUCE) continue;
      if ((j == AVF_K_PREDICT || j == AVF_K_SCAN_A) && timed[t] != 0) continue;
      if (j == AVF_K_FINALIZE && timed[t] + 1 != S) continue;
      if (j == AVF_K_COUNT && (timed[t] + 1 == S || c->clump_bins)) continue;
      const size_t i = (size_t)timed[t];
      float ms = 0.f;
      he = hipEventElapsedTime(&ms, ev[i * kEv + j], ev[i * kEv + j + 1]);
      prof->sum[j] += ms;
      prof->count[j] += 1;
    }
  if (he != hipSuccess) return fail(c, AVF_ERR_HIP, "event timing: %s", hipGetErrorString(he));
  prof->launched += (int64_t)timed.size();
  return AVF_OK;
}

// Steps with the sphere cells rebuilt every step (avf_options.verlet_skin == 0), in chunks
// that end at the spatial reorders; a chunk runs as captured graphs of at most kGraphSteps
// steps each (cached by chunk length and offset) unless profiled or use_graph is off.
int run_cells(avf_ctx *c, int64_t nsteps, StepProf *prof) {
  constexpr int64_t kGraphSteps = 64;
  const int64_t chunk_max = c->opt.reorder_interval > 0 ? c->opt.reorder_interval : 1024;
  int rc;
  int64_t left = nsteps;
  while (left > 0) {
    if (!c->sgrid_valid ||
        (c->opt.reorder_interval > 0 && c->steps_since_reorder >= c->opt.reorder_interval)) {
      if ((rc = rebuild_grid(c, c->opt.reorder_interval > 0))) return rc;
    }
    const int64_t S = std::min(left, chunk_max - (c->steps_since_reorder % chunk_max));
    if (prof) {
      if ((rc = csteps_profiled(c, S, prof))) return rc;
    } else if (c->opt.use_graph) {
      for (int64_t off = 0; off < S; off += kGraphSteps) {
        const int64_t len = std::min(kGraphSteps, S - off);
        const int64_t key = ((int64_t{1} << 50) + S * 65536 + off) ^ (c->buf_set << 56);
        auto it = c->graphs.find(key);
        if (it == c->graphs.end()) {
          hipGraph_t g;
          HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
          for (int64_t i = off; i < off + len; i++) launch_cstep(c, i, S, nullptr, i + 1 == S);
          HIPCHK(c, hipStreamEndCapture(c->stream, &g));
          hipGraphExec_t ge;
          HIPCHK(c, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
          c->graph_defs.push_back(g);
          it = c->graphs.emplace(key, ge).first;
        }
        HIPCHK(c, hipGraphLaunch(it->second, c->stream));
      }
    } else {
      for (int64_t i = 0; i < S; i++) launch_cstep(c, i, S, nullptr, i + 1 == S);
    }
    HIPCHK(c, hipGetLastError());
    c->binned = !c->clump_bins;  // the sphere arrays are the last step's with sphere keys
    c->stepped = true;
    c->steps_since_reorder += S;
    c->steps_total += S;
    left -= S;
  }
  return AVF_OK;
}

// The sphere binning of the state in st for the exports that read it, when the steps since
// it was last binned did not keep one (per-step binning by clump, candidate-list steps).
int bin_current(avf_ctx *c) {
  if (c->binned || !c->stepped || !c->sgrid_valid) return AVF_OK;
  launch_bin(c, 0);
  HIPCHK(c, hipGetLastError());
  c->binned = true;
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
  if (!(c->opt.verlet_skin > 0.0)) c->opt.verlet_skin = 0.0;
  c->clump_bins = c->opt.verlet_skin == 0.0 && c->opt.step_binning == 0;
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
  // lean chunks unless asked otherwise: the stage record holds the whole carried state but
  // the angular acceleration, so a list step reads and writes each clump's state once
  c->lean = c->opt.lean_state >= 0;
  // the steps recompute clump mass and inertia from radius and density (where they hold
  // the closed forms) unless asked otherwise: the bytes saved outweigh the arithmetic
  // (per-step binning at 250k clumps 1 % faster, candidate lists 0.5 %, at 2M clumps 2.7 %)
  // contexts of at most one wavefront per SIMD (256 CUs x 4 SIMDs x 64 clumps) are bound by
  // the latency of the candidate gathers, not by occupancy
  c->deep = total <= 256 * 4 * kForceBlock;
  c->closed = c->opt.closed_forms >= 0;
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
      (rc = dalloc(c, &c->d_nbricks, 1)) || (rc = dalloc(c, &c->d_stage2, N * kStageDoubles)) ||
      (rc = dalloc(c, &c->d_tmp64, N)) || (rc = dalloc(c, &c->d_counts, c->cell_cap)) ||
      (rc = dalloc(c, &c->d_starts, c->cell_cap + 1)) ||
      (rc = dalloc(c, &c->d_bsum, grid_for(c->cell_cap, kScanTile) + 1)) ||
      (rc = dalloc(c, &c->d_bsum2, 2 * (size_t)kScanGrid * kBsumStride)) ||
      (rc = dalloc(c, &c->d_skey, 4 * N)) || (rc = dalloc(c, &c->d_srank, 4 * N)) ||
      (rc = dalloc(c, &c->d_sid, 4 * N)) || (rc = dalloc(c, &c->d_spos, 4 * N)) ||
      (rc = dalloc(c, &c->d_clist, 4 * N * (c->opt.verlet_skin > 0.0 ? kSphListV : kSphList))) ||
      (rc = dalloc(c, &c->d_ccnt, 4 * N)) ||
      (rc = dalloc(c, &c->d_acc_cell, 8 * ninst)) || (rc = dalloc(c, &c->d_acc_sg, 8 * ninst)) ||
      (rc = dalloc(c, &c->d_diag, 1)) || (rc = dalloc(c, &c->d_count, 1)))
    return bail(rc);
  if ((he = hipHostMalloc((void **)&c->h_mail, sizeof(avf_ctx::HostMail))) != hipSuccess ||
      (he = hipHostMalloc((void **)&c->h_P, ninst * sizeof(InstParams))) != hipSuccess) {
    fail(c, AVF_ERR_NOMEM, "hipHostMalloc: %s", hipGetErrorString(he));
    return bail(AVF_ERR_NOMEM);
  }
  if ((rc = scratch_address(c))) return bail(rc);
  if ((c->opt.verlet_skin > 0.0 || c->clump_bins) &&
      ((rc = dalloc(c, &c->d_vl_m, kVMax * N)) || (rc = dalloc(c, &c->d_vl_mask, kVMax * N)) ||
       (rc = dalloc(c, &c->d_vl_cnt, N))))
    return bail(rc);
  if (c->clump_bins &&
      ((rc = dalloc(c, &c->d_ckey, N)) || (rc = dalloc(c, &c->d_crank, N)) ||
       (rc = dalloc(c, &c->d_cik, N)) || (rc = dalloc(c, &c->d_frec, N)) ||
       (rc = dalloc(c, &c->d_sca, N)) ||
       (rc = dalloc(c, &c->d_cgrid, 2))))
    return bail(rc);
  if (c->opt.verlet_skin > 0.0 &&
      ((rc = dalloc(c, &c->d_dsp, N)) || (rc = dalloc(c, &c->d_wreach, N)) ||
       (rc = dalloc(c, &c->d_ck_st, N * S_NFIELDS)) || (rc = dalloc(c, &c->d_ck_rad, N)) ||
       (rc = dalloc(c, &c->d_ck_P, ninst)) || (rc = dalloc(c, &c->d_ck_diag, 1)) ||
       (rc = dalloc(c, &c->d_umax, 1))))
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
      (he = hipMemset(c->d_bsum2, 0, 2 * (size_t)kScanGrid * kBsumStride * sizeof(uint32_t))) !=
          hipSuccess ||
      (he = hipMemset(c->d_st, 0, N * S_NFIELDS * sizeof(double))) != hipSuccess ||
      (he = hipMemset(c->d_sta, 0, N * C_NFIELDS * sizeof(double))) != hipSuccess ||
      (he = hipMemset(c->d_f, 0, N * F_NFIELDS * sizeof(double))) != hipSuccess ||
      (he = hipMemset(c->d_diag, 0, sizeof(DevDiag))) != hipSuccess ||
      (he = hipMemset(c->d_ncells, 0, sizeof(int64_t))) != hipSuccess ||
      (he = hipMemset(c->d_nbricks, 0, sizeof(int64_t))) != hipSuccess ||
      (c->d_cgrid && (he = hipMemset(c->d_cgrid, 0, 2 * sizeof(int64_t))) != hipSuccess) ||
      (c->d_cik && (he = hipMemset(c->d_cik, 0, N * sizeof(uint2))) != hipSuccess)) {
    fail(c, AVF_ERR_HIP, "hipMemset: %s", hipGetErrorString(he));
    return bail(AVF_ERR_HIP);
  }
  *out = c;
  return AVF_OK;
}

}  // extern "C"

// Buffers of the differentiable window, grown on demand and kept with the context.
struct WinBuf {
  int64_t cap = 0, nseg = 0, K = 0;
  int32_t ninst = 0, nbins = -1;
  bool tangent = false;
  int32_t *list = nullptr, *cnt = nullptr, *in = nullptr, *icnt = nullptr;
  uint8_t *eflag = nullptr;
  double *sph0 = nullptr, *dmax2 = nullptr, *wa = nullptr, *wb = nullptr, *ck = nullptr;
  double *segst = nullptr, *segstage = nullptr, *segtq = nullptr;
  double *Sb = nullptr, *Sb2 = nullptr, *sbar = nullptr, *ebuf = nullptr, *tbar = nullptr;
  double *part = nullptr, *out = nullptr, *instw = nullptr, *wbins = nullptr, *vec = nullptr;
  Dual *stD = nullptr, *stD2 = nullptr, *stageD = nullptr;
  WinParams *WP = nullptr;
  InstParams *allP = nullptr;  // parameters of every window step (clump-cell grid, cc)
  int64_t allP_steps = 0;
  long long *acc = nullptr;    // k_cellreduce accumulators of the window
  WinDiag *wd = nullptr;
  unsigned long long *umax = nullptr, *missed = nullptr;
  // the window's own binning of a state (list check)
  double *bstage = nullptr;
  double4 *bspos = nullptr;
  uint32_t *bkey = nullptr, *brank = nullptr, *bsid = nullptr, *bstarts = nullptr;
};

namespace {

void free_window(avf_ctx *c) {
  WinBuf *w = c->win;
  if (!w) return;
  void *bufs[] = {w->list,  w->cnt,   w->in,    w->icnt,     w->eflag, w->sph0,  w->dmax2,
                  w->wa,    w->wb,    w->ck,    w->segst,    w->segstage, w->segtq, w->Sb,
                  w->Sb2,   w->sbar,  w->ebuf,  w->tbar,     w->part,  w->out,   w->instw,
                  w->wbins, w->vec,   w->stD,   w->stD2,     w->stageD, w->WP,   w->wd,
                  w->umax,  w->missed, w->bstage, w->bspos, w->bkey, w->brank, w->bsid,
                  w->bstarts, w->allP, w->acc};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  delete w;
  c->win = nullptr;
}

}  // namespace

extern "C" {

int avf_destroy(avf_ctx *c) {
  if (!c) return AVF_OK;
  for (auto &kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto g : c->graph_defs) (void)hipGraphDestroy(g);
  void *bufs[] = {c->d_P,      c->d_st,    c->d_st2,   c->d_sta,   c->d_sta2,  c->d_stage,
                  c->d_f,      c->d_io,    c->d_io2,   c->d_orig,  c->d_orig2, c->d_inst,
                  c->d_inst2,  c->d_inv,   c->d_ncells, c->d_nbricks, c->d_tmp64, c->d_counts,
                  c->d_starts,
                  c->d_bsum,   c->d_bsum2, c->d_skey,  c->d_srank, c->d_sid,   c->d_spos,  c->d_acc_cell,
                  c->d_acc_sg, c->d_diag,  c->d_count, c->d_clist, c->d_ccnt,
                  c->d_vl_m,   c->d_vl_cnt, c->d_vl_mask, c->d_dsp, c->d_wreach,
                  c->d_ck_st,  c->d_ck_rad, c->d_ck_P, c->d_ck_diag, c->d_umax,
                  c->d_stage2, c->d_ckey, c->d_crank, c->d_cik, c->d_frec,
                  c->d_sca,    c->d_cgrid};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
  if (c->h_mail) (void)hipHostFree(c->h_mail);
  if (c->h_P) (void)hipHostFree(c->h_P);
  free_window(c);
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
  // mass and inertia recomputed in the step when every clump holds the closed forms
  const int32_t closed = c->closed ? 1 + c->desc[inst].geo.clump_mode : 0;
  HIPCHK(c, hipMemcpyAsync(&c->d_P[inst].mass_closed, &closed, sizeof closed,
                           hipMemcpyHostToDevice, c->stream));
  k_import_aos<<<grid_for(cnt, kBlock), kBlock, 0, c->stream>>>(
      c->d_st, c->d_sta, c->d_inv, c->begin[inst], c->d_io, c->d_io2, c->d_P + inst, cnt,
      c->cap);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->sgrid_valid = false;
  c->binned = false;
  c->stepped = false;
  c->list_valid = false;
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
  c->list_valid = false;  // a new phase, time or side.x: the next step rebuilds the lists
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
                                                                  c->d_acc_cell, c->n, c->cap, 1,
                                                                  c->ninst);
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
  c->list_valid = false;
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AVF_OK;
}

int avf_step(avf_ctx *c, int64_t nsteps) {
  if (!c || nsteps < 0) return AVF_ERR_ARG;
  const int rc = c->opt.verlet_skin > 0.0 ? run_verlet(c, nsteps, nullptr)
                                          : run_cells(c, nsteps, nullptr);
  return rc ? rc : step_status(c);
}

int avf_step_profiled(avf_ctx *c, int64_t nsteps, double *kernel_ms, double *total_ms) {
  if (!c || nsteps < 0) return AVF_ERR_ARG;
  const bool verlet = c->opt.verlet_skin > 0.0;
  hipEvent_t first, last;
  HIPCHK(c, hipEventCreate(&first));
  HIPCHK(c, hipEventCreate(&last));
  HIPCHK(c, hipEventRecord(first, c->stream));
  StepProf prof;
  int rc = verlet ? run_verlet(c, nsteps, &prof) : run_cells(c, nsteps, &prof);
  if (!rc) rc = step_status(c);
  HIPCHK(c, hipEventRecord(last, c->stream));
  HIPCHK(c, hipEventSynchronize(last));
  float tot = 0.f;
  HIPCHK(c, hipEventElapsedTime(&tot, first, last));
  (void)hipEventDestroy(first);
  (void)hipEventDestroy(last);
  // mean duration per launch of each kernel (per-step binning: over the timed launches of
  // that kernel; candidate lists: per timed step, slots of unlaunched kernels near 0)
  for (int j = 0; j < AVF_NKERNELS; j++) {
    const int64_t cnt = verlet ? prof.launched : prof.count[j];
    if (kernel_ms) kernel_ms[j] = cnt ? prof.sum[j] / (double)cnt : 0.0;
  }
  if (total_ms) *total_ms = tot;
  return rc;
}

int avf_step_timed(avf_ctx *c, int64_t nsteps, int32_t kernel, int32_t every,
                   double *kernel_ms, double *total_ms) {
  if (!c || nsteps < 0 || kernel < 0 || kernel >= AVF_NKERNELS || every < 1) return AVF_ERR_ARG;
  std::vector<double> all(AVF_NKERNELS, 0.0);
  c->prof_kernel = kernel;
  c->prof_every = every;
  c->prof_seq = 0;
  const int rc = avf_step_profiled(c, nsteps, all.data(), total_ms);
  c->prof_kernel = -1;
  c->prof_every = 1;
  if (kernel_ms) *kernel_ms = all[kernel];
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

// clear_cell / clear_ocup / cell_locate (cells.cu:11-111) on the current state, from the
// device's clump-cell indices, clumps in ascending reference index.
int avf_get_cell_table(avf_ctx *c, int32_t inst, int64_t *cell_vec, int32_t *nocup,
                       int64_t ncell3, int64_t *bad) {
  int rc = check_inst(c, inst);
  if (rc) return rc;
  if (!cell_vec || !nocup) return fail(c, AVF_ERR_ARG, "avf_get_cell_table: null table");
  InstParams p;
  HIPCHK(c, hipMemcpy(&p, c->d_P + inst, sizeof p, hipMemcpyDeviceToHost));
  const int64_t want = p.ncell[0] * p.ncell[1] * p.ncell[2];
  if (ncell3 != want)
    return fail(c, AVF_ERR_ARG, "avf_get_cell_table: ncell3 %lld, the grid has %lld cells",
                (long long)ncell3, (long long)want);
  const int64_t cnt = c->desc[inst].sys.npart, ntags = c->desc[inst].ntags;
  std::vector<int64_t> cells((size_t)cnt);
  if ((rc = avf_get_clump_cells(c, inst, cells.data()))) return rc;
  for (int64_t i = 0; i < ncell3 * ntags; i++) cell_vec[i] = -1;
  for (int64_t i = 0; i < ncell3; i++) nocup[i] = 0;
  if (bad) *bad = -1;
  for (int64_t m = 0; m < cnt; m++) {
    const int64_t idx = cells[(size_t)m];
    if (idx == -1) continue;  // frozen (cells.cu:65)
    if (idx == -2 || idx >= ncell3) {  // exits, or writes past the table (idx > ncell)
      if (bad) *bad = m;
      return fail(c, AVF_ERR_CELL_NEG, "clump %lld outside the clump-cell grid", (long long)m);
    }
    if (nocup[idx] == ntags - 1) {
      if (bad) *bad = m;
      return fail(c, AVF_ERR_CELL_FULL, "clump cell %lld full at clump %lld", (long long)idx,
                  (long long)m);
    }
    cell_vec[idx * ntags + nocup[idx]] = m;
    nocup[idx]++;
  }
  return AVF_OK;
}

// Sphere keys of the last binning, instance-relative, and the order of the counting sort.
int avf_get_sphere_keys(avf_ctx *c, int32_t inst, uint32_t *keys, int64_t *order,
                        int64_t *nbinned, avf_sphere_grid *grid) {
  int rc = check_inst(c, inst);
  if (rc) return rc;
  if ((rc = bin_current(c))) return rc;
  if (!c->binned)
    return fail(c, AVF_ERR_ARG,
                "no sphere binning of the current state (call avf_initial_forces or step)");
  InstParams p;
  HIPCHK(c, hipMemcpy(&p, c->d_P + inst, sizeof p, hipMemcpyDeviceToHost));
  if (grid) {
    for (int k = 0; k < 3; k++) {
      grid->lo[k] = p.sg_lo[k];
      grid->n[k] = p.sg_n[k];
    }
    grid->h = p.sg_h;
    grid->inv_h = p.sg_inv_h;
    grid->reserved = 0;
  }
  const int64_t cnt = c->desc[inst].sys.npart, b = c->begin[inst];
  std::vector<uint32_t> sk((size_t)(4 * cnt));
  std::vector<int32_t> orig((size_t)cnt);
  HIPCHK(c, hipMemcpy(sk.data(), c->d_skey + 4 * b, sk.size() * sizeof(uint32_t),
                      hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(orig.data(), c->d_orig + b, orig.size() * sizeof(int32_t),
                      hipMemcpyDeviceToHost));
  std::vector<uint32_t> kref((size_t)(4 * cnt), 0xffffffffu);
  for (int64_t k = 0; k < cnt; k++) {
    const int64_t o = orig[(size_t)k];
    if (o < 0 || o >= cnt) return fail(c, AVF_ERR_HIP, "clump order table corrupt");
    for (int s = 0; s < 4; s++) {
      const uint32_t key = sk[(size_t)(4 * k + s)];
      kref[(size_t)(4 * o + s)] = key == kInvalidKey ? 0xffffffffu : key - (uint32_t)p.sg_base;
    }
  }
  if (keys) std::memcpy(keys, kref.data(), kref.size() * sizeof(uint32_t));
  std::vector<int64_t> ord;
  ord.reserve(kref.size());
  for (int64_t i = 0; i < 4 * cnt; i++)
    if (kref[(size_t)i] != 0xffffffffu) ord.push_back(i);
  std::stable_sort(ord.begin(), ord.end(),
                   [&kref](int64_t a, int64_t b2) { return kref[(size_t)a] < kref[(size_t)b2]; });
  if (order) std::memcpy(order, ord.data(), ord.size() * sizeof(int64_t));
  if (nbinned) *nbinned = (int64_t)ord.size();
  return AVF_OK;
}

int avf_get_contacts(avf_ctx *c, int32_t inst, avf_contact *out, int64_t capn, int64_t *count) {
  int rc = check_inst(c, inst);
  if (rc) return rc;
  if ((rc = bin_current(c))) return rc;
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

}  // extern "C"

namespace {

// Observable sums of one instance for the state st (kObsScalars + nbins doubles; see
// k_observe) into h.
int observe_state(avf_ctx *c, const double *st, int32_t inst, int nbins, double x0, double bw,
                  double sigma, double width, std::vector<double> &h) {
  const int64_t cnt = c->desc[inst].sys.npart;
  const int64_t nb = (cnt + kObsBlock - 1) / kObsBlock;
  const int64_t w = kObsScalars + nbins;
  double *d_part = nullptr, *d_out = nullptr;
  int rc;
  if ((rc = dalloc(c, &d_part, nb * w)) || (rc = dalloc(c, &d_out, w))) {
    if (d_part) (void)hipFree(d_part);
    return rc;
  }
  k_observe<<<(unsigned)nb, kObsBlock, 0, c->stream>>>(st, c->d_sta, c->d_inv, c->d_P, inst, cnt,
                                                      c->cap, nbins, x0, bw, sigma, width, d_part);
  k_observe_final<<<1, 1, 0, c->stream>>>(d_part, nb, nbins, d_out);
  h.assign((size_t)w, 0.0);
  hipError_t he = hipGetLastError();
  if (he == hipSuccess)
    he = hipMemcpyAsync(h.data(), d_out, w * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
  (void)hipFree(d_part);
  (void)hipFree(d_out);
  if (he != hipSuccess) return fail(c, AVF_ERR_HIP, "observe: %s", hipGetErrorString(he));
  return AVF_OK;
}

int profile_params(avf_ctx *c, const avf_profile_spec *spec, int *nbins, double *x0, double *bw,
                   double *sigma, double *width) {
  *nbins = 0;
  *x0 = 0.0;
  *bw = *sigma = *width = 1.0;
  if (!spec || spec->nbins == 0) return AVF_OK;
  if (spec->nbins < 0 || spec->nbins > kMaxBins || !(spec->x1 > spec->x0) || !(spec->width > 0.0))
    return fail(c, AVF_ERR_ARG, "bad deposit profile spec (nbins 1..%d, x1 > x0, width > 0)",
                kMaxBins);
  *nbins = spec->nbins;
  *x0 = spec->x0;
  *bw = (spec->x1 - spec->x0) / spec->nbins;
  *sigma = spec->sigma > 0.0 ? spec->sigma : *bw;
  *width = spec->width;
  return AVF_OK;
}

void fill_observables(avf_ctx *c, int32_t inst, const double *h, avf_observables *obs) {
  const int64_t cnt = c->desc[inst].sys.npart;
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

}  // namespace

extern "C" {

int avf_observe(avf_ctx *c, int32_t inst, avf_observables *obs, const avf_profile_spec *spec,
                double *profile) {
  int rc = check_inst(c, inst);
  if (rc) return rc;
  int nbins;
  double x0, bw, sigma, width;
  if ((rc = profile_params(c, profile ? spec : nullptr, &nbins, &x0, &bw, &sigma, &width)))
    return rc;
  std::vector<double> h;
  if ((rc = observe_state(c, c->d_st, inst, nbins, x0, bw, sigma, width, h))) return rc;
  if (obs) fill_observables(c, inst, h.data(), obs);
  for (int b = 0; b < nbins; b++) profile[b] = h[kObsScalars + b];
  return AVF_OK;
}

int avf_observe_noalloc(avf_ctx *c, int32_t inst, avf_observables *obs,
                        const avf_profile_spec *spec, double *profile) {
  int rc = check_inst(c, inst);
  if (rc) return rc;
  int nbins;
  double x0, bw, sigma, width;
  if ((rc = profile_params(c, profile ? spec : nullptr, &nbins, &x0, &bw, &sigma, &width)))
    return rc;
  // the same kernels and grid as avf_observe, the block partials and the sums in scratch
  const int64_t cnt = c->desc[inst].sys.npart;
  const int64_t nb = (cnt + kObsBlock - 1) / kObsBlock;
  const int64_t w = kObsScalars + nbins;
  if ((nb + 1) * w > kObsScratchDoubles)
    return fail(c, AVF_ERR_NOMEM,
                "avf_observe_noalloc: %lld clumps with %d bins need %lld doubles of scratch, "
                "more than its %lld", (long long)cnt, nbins, (long long)((nb + 1) * w),
                (long long)kObsScratchDoubles);
  double *part = c->d_scratch, *sums = c->d_scratch + nb * w;
  k_observe<<<(unsigned)nb, kObsBlock, 0, c->stream>>>(c->d_st, c->d_sta, c->d_inv, c->d_P, inst,
                                                      cnt, c->cap, nbins, x0, bw, sigma, width,
                                                      part);
  k_observe_final<<<1, 1, 0, c->stream>>>(part, nb, nbins, sums);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(c->h_mail->obs, sums, w * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (obs) fill_observables(c, inst, c->h_mail->obs, obs);
  for (int b = 0; b < nbins; b++) profile[b] = c->h_mail->obs[kObsScalars + b];
  return AVF_OK;
}

int avf_get_diag(avf_ctx *c, avf_diag *d) {
  if (!c || !d) return AVF_ERR_ARG;
  HIPCHK(c, hipMemcpyAsync(&c->h_mail->diag, c->d_diag, sizeof(DevDiag), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_diag, 0, sizeof(DevDiag), c->stream));
  HIPCHK(c, hipMemcpyAsync(c->h_P, c->d_P, c->ninst * sizeof(InstParams), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const DevDiag h = c->h_mail->diag;
  c->seen_below = c->seen_overflow = c->seen_broken = 0;  // the device counters restart from zero
  int64_t closed = 0;
  for (int32_t i = 0; i < c->ninst; i++) closed += c->h_P[i].mass_closed != 0;
  d->cell_below = (int64_t)h.cell_below;
  d->cell_beyond = (int64_t)h.cell_beyond;
  d->neigh_overflow = (int64_t)h.neigh_overflow;
  d->max_neighbours = (int64_t)h.max_neighbours;
  d->max_cell_spheres = (int64_t)h.max_cell;
  d->sweep_global_bricks = (int64_t)h.sweep_global;
  d->sweep_fp64_pairs = (int64_t)h.sweep_fp64;
  d->steps = c->steps_total;
  d->list_builds = c->list_builds;
  d->steps_undone = c->steps_undone;
  d->closed_instances = closed;
  return AVF_OK;
}

}  // extern "C"

// ------------------------------------------------------------ differentiable window ----

namespace {

constexpr double kPiW = 3.14159265358979323846;

// gamma(eps, kappa) = -2 ln(eps) sqrt(kappa / (pi^2 + ln^2 eps)) and its partials; with
// t_c = sqrt((pi^2 + ln^2 eps) / kappa) this is the reference's gamma = -2 ln(eps) / t_c.
struct LawMap {
  double kappa, gamma, mu, fr, dg_deps, dg_dkappa;
};

LawMap law_from_theta(double mu, double eps, double kappa) {
  LawMap m;
  const double L = std::log(eps), S2 = kPiW * kPiW + L * L, S = std::sqrt(S2);
  m.kappa = kappa;
  m.gamma = -2.0 * L * std::sqrt(kappa / S2);
  m.mu = mu;
  m.fr = mu / 1.0e12;
  m.dg_deps = -2.0 * std::sqrt(kappa) * kPiW * kPiW / (eps * S2 * S);
  m.dg_dkappa = m.gamma / (2.0 * kappa);
  return m;
}

// The instance's own law (values exactly as derived by the setup), with the partials of
// the same relation.
LawMap law_from_system(double mu, double eps, double kappa, double gamma, double fr) {
  LawMap m = law_from_theta(mu, eps, kappa);
  m.gamma = gamma;
  m.fr = fr;
  m.dg_dkappa = gamma / (2.0 * kappa);
  return m;
}

template <typename T>
int walloc(avf_ctx *c, T **p, int64_t count) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  return dalloc(c, p, count);
}

int ensure_window(avf_ctx *c, int64_t W, int64_t nseg, int64_t K, bool tangent) {
  if (!c->win) c->win = new WinBuf();
  WinBuf *w = c->win;
  const int64_t cap = c->cap;
  int rc;
  if (w->cap != cap) {
    const int64_t nb = (cap + kObsBlock - 1) / kObsBlock;
    if ((rc = walloc(c, &w->list, cap * kWinMaxList)) || (rc = walloc(c, &w->cnt, cap)) ||
        (rc = walloc(c, &w->in, cap * kWinMaxIn)) || (rc = walloc(c, &w->icnt, cap)) ||
        (rc = walloc(c, &w->eflag, cap * kWinMaxList)) || (rc = walloc(c, &w->sph0, cap * 12)) ||
        (rc = walloc(c, &w->dmax2, cap)) || (rc = walloc(c, &w->wa, cap * S_NFIELDS)) ||
        (rc = walloc(c, &w->wb, cap * S_NFIELDS)) || (rc = walloc(c, &w->Sb, cap * S_NFIELDS)) ||
        (rc = walloc(c, &w->Sb2, cap * S_NFIELDS)) || (rc = walloc(c, &w->sbar, cap * kStageW)) ||
        (rc = walloc(c, &w->ebuf, cap * kWinMaxList * kEbuf)) ||
        (rc = walloc(c, &w->tbar, cap * TB_N)) || (rc = walloc(c, &w->part, nb * TB_N)) ||
        (rc = walloc(c, &w->out, (int64_t)c->ninst * TB_N)) ||
        (rc = walloc(c, &w->vec, cap * 6)) || (rc = walloc(c, &w->WP, c->ninst)) ||
        (rc = walloc(c, &w->wd, 1)) || (rc = walloc(c, &w->umax, 1)) ||
        (rc = walloc(c, &w->missed, 1)) || (rc = walloc(c, &w->bstage, cap * kStageDoubles)) ||
        (rc = walloc(c, &w->bspos, cap * 4)) || (rc = walloc(c, &w->bkey, cap * 4)) ||
        (rc = walloc(c, &w->brank, cap * 4)) || (rc = walloc(c, &w->bsid, cap * 4)) ||
        (rc = walloc(c, &w->bstarts, c->cell_cap + 1)) ||
        (rc = walloc(c, &w->acc, (int64_t)8 * c->ninst)) ||
        (rc = walloc(c, &w->instw, c->ninst)) ||
        (rc = walloc(c, &w->wbins, (int64_t)c->ninst * kMaxBins)))
      return rc;
    w->cap = cap;
    w->nseg = w->K = w->allP_steps = 0;
    w->tangent = false;
  }
  if (W + 1 > w->allP_steps) {
    if ((rc = walloc(c, &w->allP, (W + 1) * c->ninst))) return rc;
    w->allP_steps = W + 1;
  }
  if (nseg > w->nseg) {
    if ((rc = walloc(c, &w->ck, nseg * cap * S_NFIELDS))) return rc;
    w->nseg = nseg;
  }
  if (K > w->K) {
    if ((rc = walloc(c, &w->segst, K * cap * S_NFIELDS)) ||
        (rc = walloc(c, &w->segstage, K * cap * kStageW)) ||
        (rc = walloc(c, &w->segtq, K * cap * 3)))
      return rc;
    w->K = K;
  }
  if (tangent && !w->tangent) {
    if ((rc = walloc(c, &w->stD, cap * S_NFIELDS)) || (rc = walloc(c, &w->stD2, cap * S_NFIELDS)) ||
        (rc = walloc(c, &w->stageD, cap * kStageW)))
      return rc;
    w->tangent = true;
  }
  return AVF_OK;
}

// One window step (double) with the step's parameters P: stage into stg, new state into
// out.
void win_step(avf_ctx *c, const InstParams *P, const double *cur, double *out, double *stg,
              double *tq, const double *sph0, double *dmax2, WinDiag *wd) {
  WinBuf *w = c->win;
  k_win_predict<double><<<grid_for(c->n, kBlock), kBlock, 0, c->stream>>>(cur, c->d_sta, P,
                                                                         c->d_inst, stg, c->n,
                                                                         c->cap);
  k_win_force<double><<<grid_for(c->n, 64), 64, 0, c->stream>>>(
      cur, out, c->d_sta, stg, P, w->WP, c->d_inst, w->list, w->cnt, tq, sph0, dmax2, wd, c->n,
      c->cap);
}

// Bins the state st into the window's own sphere arrays (stage, sorted spheres); the
// context's own binning stays as it is.
void win_bin(avf_ctx *c, const double *st) {
  WinBuf *w = c->win;
  k_predict<<<grid_for(c->n, kPredictBlock), kPredictBlock, 0, c->stream>>>(
      const_cast<double *>(st), c->d_sta, w->bstage, c->d_inst, c->d_orig, c->d_P, c->d_counts,
      w->bkey, w->brank, c->n, c->cap, 0, ListTrack{nullptr, nullptr}, 0, nullptr);
  launch_scan(c, w->bstarts);
  launch_scatter(c, w->bkey, w->brank, w->bstarts, w->bstage, w->bspos, w->bsid, nullptr);
}

// Adds the touching sphere pairs of state st that the frozen list lacks to w->missed.
void win_check_list(avf_ctx *c, const double *st) {
  WinBuf *w = c->win;
  win_bin(c, st);
  k_win_missed<<<grid_for(4 * c->n, kBlock), kBlock, 0, c->stream>>>(
      w->bstage, c->d_P, c->d_inst, w->bstarts, w->bspos, w->bsid, w->list, w->cnt, w->missed,
      4 * c->n);
}

// Per-instance sums, in reference order, of a per-clump AoS array of width wdt (vals) or of
// the tangent dot product (vals NULL) -> host out[inst * wdt + k].
int win_reduce(avf_ctx *c, const double *vals, int wdt, const Dual *stD, std::vector<double> &out) {
  WinBuf *w = c->win;
  out.assign((size_t)c->ninst * wdt, 0.0);
  for (int32_t i = 0; i < c->ninst; i++) {
    const int64_t cnt = c->desc[i].sys.npart;
    const int64_t nb = (cnt + kObsBlock - 1) / kObsBlock;
    k_win_partial<<<(unsigned)nb, kObsBlock, 0, c->stream>>>(vals, wdt, w->Sb, stD, c->d_inv,
                                                            c->begin[i], cnt, c->cap, w->part);
    k_win_partial_final<<<1, 64, 0, c->stream>>>(w->part, nb, wdt, w->out + (int64_t)i * wdt);
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(out.data(), w->out, out.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AVF_OK;
}

int read_wdiag(avf_ctx *c, WinDiag *h) {
  HIPCHK(c, hipMemcpyAsync(h, c->win->wd, sizeof *h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AVF_OK;
}

}  // namespace

extern "C" int avf_window_run(avf_ctx *c, const avf_window_spec *spec, const double *theta,
                              const double *dv, const double *target, double *loss,
                              avf_observables *obs, double *profile, double *grad_theta,
                              double *grad_dv, const double *tan_theta, const double *tan_dv,
                              double *dloss, double *final_state, avf_window_diag *diag) {
  if (!c || !spec || !loss) return AVF_ERR_ARG;
  const int64_t W = spec->steps;
  const int64_t K = spec->checkpoint > 0 ? std::min(spec->checkpoint, W) : W;
  if (W <= 0) return fail(c, AVF_ERR_ARG, "window length must be positive");
  if (!(spec->skin >= 0.0)) return fail(c, AVF_ERR_ARG, "skin must be >= 0");
  if (c->cap * kWinMaxList > (int64_t)INT32_MAX)
    return fail(c, AVF_ERR_ARG, "window supports up to %lld clumps per context",
                (long long)(INT32_MAX / kWinMaxList));
  for (int32_t i = 0; i < c->ninst; i++) {
    if (c->desc[i].phase != AVF_PHASE_AVALANCHE)
      return fail(c, AVF_ERR_INPUT, "instance %d: the window needs the avalanche phase", i);
    if (c->desc[i].geo.modules_avalanche & (AVF_MOD_FUNNEL | AVF_MOD_FCORNERS))
      return fail(c, AVF_ERR_INPUT,
                  "instance %d: the window does not support the funnel / funnel-corner modules", i);
  }
  int rc, nbins;
  double x0, bw, sigma, width;
  if ((rc = profile_params(c, &spec->profile, &nbins, &x0, &bw, &sigma, &width))) return rc;
  const int64_t nseg = (W + K - 1) / K;
  const bool want_tan = dloss != nullptr;
  const bool want_grad = grad_theta != nullptr || grad_dv != nullptr;
  if ((rc = ensure_window(c, W, nseg, K, want_tan))) return rc;
  WinBuf *w = c->win;
  const int64_t n = c->n, cap = c->cap;

  // contact-law parameters and tangent directions
  std::vector<WinParams> hp((size_t)c->ninst);
  std::vector<LawMap> maps((size_t)c->ninst * 2);
  for (int32_t i = 0; i < c->ninst; i++) {
    const avf_system &s = c->desc[i].sys;
    for (int set = 0; set < 2; set++) {
      LawMap m;
      if (theta) {
        const double *t = theta + (size_t)i * AVF_NTHETA + 3 * set;
        if (!(t[1] > 0.0 && t[1] < 1.0) || !(t[2] > 0.0) || !(t[0] >= 0.0))
          return fail(c, AVF_ERR_ARG, "instance %d: theta needs mu >= 0, 0 < eps < 1, kappa > 0",
                      i);
        m = law_from_theta(t[0], t[1], t[2]);
      } else if (set == 0) {
        m = law_from_system(s.mu_pp, s.eps_pp, s.kappa_pp, s.gamma_pp, s.fr_fact_pp);
      } else {
        m = law_from_system(s.mu_pw, s.eps_pw, s.kappa_pw, s.gamma_pw, s.fr_fact_pw);
      }
      maps[(size_t)i * 2 + set] = m;
      WinParams &wp = hp[(size_t)i];
      wp.v[set][WP_KAPPA] = m.kappa;
      wp.v[set][WP_GAMMA] = m.gamma;
      wp.v[set][WP_MU] = m.mu;
      wp.v[set][WP_FR] = m.fr;
      double tm = 0.0, te = 0.0, tk = 0.0;
      if (tan_theta) {
        const double *t = tan_theta + (size_t)i * AVF_NTHETA + 3 * set;
        tm = t[0];
        te = t[1];
        tk = t[2];
      }
      wp.d[set][WP_KAPPA] = tk;
      wp.d[set][WP_GAMMA] = m.dg_deps * te + m.dg_dkappa * tk;
      wp.d[set][WP_MU] = tm;
      wp.d[set][WP_FR] = tm / 1.0e12;
    }
  }
  hipEvent_t ev[4];
  for (auto &e : ev) HIPCHK(c, hipEventCreate(&e));
  struct EvGuard {
    hipEvent_t *e;
    ~EvGuard() {
      for (int k = 0; k < 4; k++) (void)hipEventDestroy(e[k]);
    }
  } guard{ev};
  HIPCHK(c, hipEventRecord(ev[0], c->stream));
  HIPCHK(c, hipMemcpyAsync(w->WP, hp.data(), hp.size() * sizeof(WinParams), hipMemcpyHostToDevice,
                           c->stream));
  // sphere grid of the current positions (kept as it is when valid; the storage order is not
  // changed), the window's own binning of the start state, the frozen list and its transpose
  if (!c->sgrid_valid) {
    if ((rc = rebuild_grid(c, false))) return rc;
    c->binned = false;
  }
  win_bin(c, c->d_st);
  HIPCHK(c, hipMemsetAsync(w->wd, 0, sizeof(WinDiag), c->stream));
  HIPCHK(c, hipMemsetAsync(w->icnt, 0, cap * sizeof(int32_t), c->stream));
  HIPCHK(c, hipMemsetAsync(w->umax, 0, sizeof(unsigned long long), c->stream));
  HIPCHK(c, hipMemsetAsync(w->missed, 0, sizeof(unsigned long long), c->stream));
  k_win_list<<<grid_for(n, 64), 64, 0, c->stream>>>(
      c->d_st, c->d_sta, w->bstage, c->d_P, c->d_inst, c->d_orig, w->bstarts, w->bspos, w->bsid,
      spec->skin, w->list, w->cnt, w->sph0, w->dmax2, w->wd, n, cap);
  k_win_in_fill<<<grid_for(n, kBlock), kBlock, 0, c->stream>>>(w->list, w->cnt, w->in, w->icnt,
                                                               w->wd, n);
  k_win_in_sort<<<grid_for(n, kBlock), kBlock, 0, c->stream>>>(w->in, w->icnt, c->d_orig,
                                                               w->wd, n);
  // start state (checkpoint 0) with the velocity perturbation, and its dual lift
  if (dv)
    HIPCHK(c, hipMemcpyAsync(w->vec, dv, n * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (want_tan && tan_dv)
    HIPCHK(c, hipMemcpyAsync(w->vec + 3 * cap, tan_dv, n * 3 * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
  k_win_start<<<grid_for(n, kBlock), kBlock, 0, c->stream>>>(
      c->d_st, w->ck, c->d_inv, dv ? w->vec : nullptr, (want_tan && tan_dv) ? w->vec + 3 * cap : nullptr,
      want_tan ? w->stD : nullptr, n, cap);
  HIPCHK(c, hipGetLastError());
  WinDiag hd;
  if ((rc = read_wdiag(c, &hd))) return rc;
  if (hd.frozen)
    return fail(c, AVF_ERR_INPUT,
                "%llu clumps are past side.x or before the gate opening (time <= 0); the window "
                "needs live clumps in the avalanche phase", hd.frozen);
  if (hd.aniso)
    return fail(c, AVF_ERR_INPUT, "%llu clumps have anisotropic inertia; the window needs "
                "isotropic clumps (clump_mode 1)", hd.aniso);
  if (hd.list_overflow || hd.in_overflow)
    return fail(c, AVF_ERR_NEIGH_OVERFLOW,
                "window list overflow (%llu clumps beyond %d pairs / %d neighbours; %llu beyond "
                "%d incoming entries): reduce the skin", hd.list_overflow, kWinMaxList,
                kWinMaxNeigh, hd.in_overflow, kWinMaxIn);

  // forward window with checkpoints at the segment starts; the parameters of every step
  // (the clump-cell grid is regridded on the plain step's schedule) are kept
  const size_t SF = (size_t)cap * S_NFIELDS;
  const int32_t ni = c->ninst;
  {
    std::vector<long long> acc0;
    reset_acc(acc0, ni);
    HIPCHK(c, hipMemcpyAsync(w->acc, acc0.data(), acc0.size() * sizeof(long long),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(w->allP, c->d_P, ni * sizeof(InstParams), hipMemcpyDeviceToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  const double *cur = w->ck;
  for (int64_t t = 0; t < W; t++) {
    double *out;
    if ((t + 1) % K == 0 && t + 1 < W)
      out = w->ck + (size_t)((t + 1) / K) * SF;
    else
      out = (cur == w->wa) ? w->wb : w->wa;
    InstParams *Pt = w->allP + (size_t)t * ni;
    win_step(c, Pt, cur, out, w->segstage, nullptr, w->sph0, w->dmax2, w->wd);
    k_cellreduce<<<grid_for(n, kBlock), kBlock, 0, c->stream>>>(out, c->d_inst, Pt, w->acc, n, cap,
                                                                0, ni);
    k_win_finalize<<<grid_for(ni, 64), 64, 0, c->stream>>>(Pt, Pt + ni, w->acc, ni);
    cur = out;
    if ((t + 1) % K == 0 || t + 1 == W) win_check_list(c, cur);
  }
  HIPCHK(c, hipGetLastError());
  const double *fin = cur;
  if ((rc = read_wdiag(c, &hd))) return rc;
  if (hd.left)
    return fail(c, AVF_ERR_INPUT, "a clump reached x > side.x within the window (%llu clump-steps)",
                hd.left);
  // loss of the final state, and the adjoint seed
  std::vector<double> instw((size_t)c->ninst), wbins((size_t)c->ninst * std::max(nbins, 1), 0.0);
  for (int32_t i = 0; i < c->ninst; i++) {
    std::vector<double> h;
    if ((rc = observe_state(c, fin, i, nbins, x0, bw, sigma, width, h))) return rc;
    avf_observables o;
    fill_observables(c, i, h.data(), &o);
    double L = spec->w_runout * o.runout_mean;
    for (int b = 0; b < nbins; b++) {
      const double tb = target ? target[(size_t)i * nbins + b] : 0.0;
      const double r = h[kObsScalars + b] - tb;
      L += spec->w_profile * r * r;
      wbins[(size_t)i * nbins + b] = 2.0 * spec->w_profile * r;
    }
    loss[i] = L;
    instw[i] = o.mass_live > 0.0 ? spec->w_runout / o.mass_live : 0.0;
    if (obs) obs[i] = o;
    if (profile)
      for (int b = 0; b < nbins; b++) profile[(size_t)i * nbins + b] = h[kObsScalars + b];
  }
  if (final_state) {
    double *d_fs = nullptr;
    if ((rc = dalloc(c, &d_fs, n * S_NFIELDS))) return rc;
    k_win_export<<<grid_for(n, kBlock), kBlock, 0, c->stream>>>(fin, nullptr, c->d_inv, d_fs,
                                                                nullptr, n, cap);
    hipError_t he = hipGetLastError();
    if (he == hipSuccess)
      he = hipMemcpyAsync(final_state, d_fs, n * S_NFIELDS * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    (void)hipFree(d_fs);
    if (he != hipSuccess) return fail(c, AVF_ERR_HIP, "window export: %s", hipGetErrorString(he));
  }
  HIPCHK(c, hipEventRecord(ev[1], c->stream));
  HIPCHK(c, hipMemcpyAsync(w->instw, instw.data(), instw.size() * sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(w->wbins, wbins.data(), wbins.size() * sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
  k_win_loss_grad<<<grid_for(n, kBlock), kBlock, 0, c->stream>>>(
      fin, c->d_sta, c->d_P, c->d_inst, w->instw, w->wbins, nbins,
      x0, bw, sigma, width, w->Sb, n, cap);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));  // host vectors above stay alive until here

  // dual-number forward along (tan_theta, tan_dv): dloss = seed . d(final state)
  float ms_tan = 0.f;
  if (want_tan) {
    hipEvent_t t0 = ev[2];
    HIPCHK(c, hipEventRecord(t0, c->stream));
    const Dual *dcur = w->stD;
    for (int64_t t = 0; t < W; t++) {
      Dual *dout = (dcur == w->stD) ? w->stD2 : w->stD;
      const InstParams *Pt = w->allP + (size_t)t * ni;
      k_win_predict<Dual><<<grid_for(n, kBlock), kBlock, 0, c->stream>>>(
          dcur, c->d_sta, Pt, c->d_inst, w->stageD, n, cap);
      k_win_force<Dual><<<grid_for(n, 64), 64, 0, c->stream>>>(
          dcur, dout, c->d_sta, w->stageD, Pt, w->WP, c->d_inst, w->list, w->cnt, nullptr,
          nullptr, nullptr, nullptr, n, cap);
      dcur = dout;
    }
    HIPCHK(c, hipGetLastError());
    std::vector<double> d;
    if ((rc = win_reduce(c, nullptr, 1, dcur, d))) return rc;
    for (int32_t i = 0; i < c->ninst; i++) dloss[i] = d[(size_t)i];
    HIPCHK(c, hipEventRecord(ev[3], c->stream));
    HIPCHK(c, hipEventSynchronize(ev[3]));
    HIPCHK(c, hipEventElapsedTime(&ms_tan, t0, ev[3]));
  }

  // reverse pass: per segment, recompute the steps from the checkpoint (keeping each step's
  // stage and torque), then reverse them
  float ms_bwd = 0.f;
  if (want_grad) {
    hipEvent_t b0 = ev[2], b1 = ev[3];
    HIPCHK(c, hipEventRecord(b0, c->stream));
    HIPCHK(c, hipMemsetAsync(w->tbar, 0, (size_t)cap * TB_N * sizeof(double), c->stream));
    const size_t SG = (size_t)cap * kStageW, TQ = (size_t)cap * 3;
    double *Sb = w->Sb, *Sb2 = w->Sb2;
    for (int64_t s = nseg - 1; s >= 0; s--) {
      const int64_t L = std::min(K, W - s * K);
      const double *src = w->ck + (size_t)s * SF;
      for (int64_t k = 0; k < L; k++) {
        double *dst = w->segst + (size_t)k * SF;
        win_step(c, w->allP + (size_t)(s * K + k) * ni, src, dst, w->segstage + (size_t)k * SG,
                 w->segtq + (size_t)k * TQ, nullptr, nullptr, nullptr);
        src = dst;
      }
      for (int64_t k = L - 1; k >= 0; k--) {
        const double *pre = k == 0 ? w->ck + (size_t)s * SF : w->segst + (size_t)(k - 1) * SF;
        k_win_adj_force<<<grid_for(n, 64), 64, 0, c->stream>>>(
            Sb, Sb2, c->d_sta, w->segstage + (size_t)k * SG, w->segtq + (size_t)k * TQ,
            w->allP + (size_t)(s * K + k) * ni, w->WP, c->d_inst, w->list, w->cnt, w->sbar,
            w->ebuf, w->eflag, w->tbar, n, cap);
        k_win_adj_gather<<<grid_for(n, 64), 64, 0, c->stream>>>(
            pre, Sb2, c->d_sta, w->sbar, w->ebuf, w->eflag, w->list, w->in, w->icnt, c->d_P,
            c->d_inst, n, cap);
        std::swap(Sb, Sb2);
      }
    }
    HIPCHK(c, hipGetLastError());
    if (Sb != w->Sb) {
      // keep the start-state adjoint in w->Sb (win_reduce's tangent path reads w->Sb)
      HIPCHK(c, hipMemcpyAsync(w->Sb, Sb, SF * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    }
    if (grad_dv) {
      k_win_export<<<grid_for(n, kBlock), kBlock, 0, c->stream>>>(nullptr, w->Sb, c->d_inv,
                                                                  nullptr, w->vec, n, cap);
      HIPCHK(c, hipGetLastError());
      HIPCHK(c, hipMemcpyAsync(grad_dv, w->vec, n * 3 * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
    }
    std::vector<double> tb;
    if ((rc = win_reduce(c, w->tbar, TB_N, nullptr, tb))) return rc;
    if (grad_theta) {
      for (int32_t i = 0; i < c->ninst; i++) {
        const double *t = &tb[(size_t)i * TB_N];
        for (int set = 0; set < 2; set++) {
          const LawMap &m = maps[(size_t)i * 2 + set];
          const double Kb = t[set ? TB_KPW : TB_KPP], Gb = t[set ? TB_GPW : TB_GPP];
          const double Mb = t[set ? TB_MPW : TB_MPP];
          double *g = grad_theta + (size_t)i * AVF_NTHETA + 3 * set;
          g[0] = Mb;
          g[1] = Gb * m.dg_deps;
          g[2] = Kb + Gb * m.dg_dkappa;
        }
      }
    }
    HIPCHK(c, hipEventRecord(b1, c->stream));
    HIPCHK(c, hipEventSynchronize(b1));
    HIPCHK(c, hipEventElapsedTime(&ms_bwd, b0, b1));
  }
  if (diag) {
    k_win_max<<<std::min(kWinMaxGrid, grid_for(n, kWinMaxBlock)), kWinMaxBlock, 0,
                c->stream>>>(w->dmax2, n, w->umax);
    unsigned long long um = 0, miss = 0;
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(&um, w->umax, sizeof um, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&miss, w->missed, sizeof miss, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double d2;
    std::memcpy(&d2, &um, sizeof d2);
    float ms_fwd = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms_fwd, ev[0], ev[1]));
    diag->max_disp = std::sqrt(d2);
    diag->max_list = (int64_t)hd.max_list;
    diag->max_in = (int64_t)hd.max_in;
    diag->missed = (int64_t)miss;
    diag->list_valid = miss == 0 ? 1 : 0;
    diag->reserved = 0;
    diag->forward_ms = ms_fwd;
    diag->tangent_ms = ms_tan;
    diag->backward_ms = ms_bwd;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AVF_OK;
}
