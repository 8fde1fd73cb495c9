// dem_kernels.hip -- gfx950 kernels of the DEM step for tetrahedral 4-sphere clumps.
//
// One step with the cell list rebuilt every step, binned by clump (reference:
// avalanche_hst.cu:668-710, avalanche; :438-467, hopper filling):
//   k_count_scan  exclusive scan of the clump-cell counts (the step's time and counter);
//   k_cscatter    each clump's centre and radius, storage index and key into cell order;
//   k_csweep      the candidate pairings of every clump from its 27 clump cells (a float
//                 test with a margin: a superset of the touching ones);
//   k_force       boundary modules + force_part on the listed pairings (re-tested in FP64,
//                 contacts folded in the reference's summation order) + get_accel +
//                 finish_Verlet + the cell_dim reduction when the regrid is due, then the
//                 next step's init_Verlet.cu:11-152 predict, 176-B stage record, clump key
//                 and count (integer atomic ranks: counting sort) and float record;
//   k_finalize    the regrid arithmetic and cc++ at the end of a chunk.
// A chunk's first step starts with k_predict and k_count_sum.  Binned by sphere instead,
// k_scatter and k_sweep take the place of k_cscatter and k_csweep; on candidate-list steps
// k_force alone makes the step (k_vmerge builds the lists).  No float atomics anywhere:
// per-clump sums are sequential in the reference order.
#include "dem_kernels.h"

namespace avf {

// ----------------------------------------------------------------------- helpers ----

__device__ __forceinline__ int grid_index(double x, double lo, double inv_h, int n) {
  double f = (x - lo) * inv_h;
  if (!(f >= 0.0)) return 0;
  if (f >= (double)n) return n - 1;
  return (int)f;
}

__device__ __forceinline__ uint32_t sphere_key(const InstParams &p, d3 q) {
  int ix = grid_index(q.x, p.sg_lo[0], p.sg_inv_h, p.sg_n[0]);
  int iy = grid_index(q.y, p.sg_lo[1], p.sg_inv_h, p.sg_n[1]);
  int iz = grid_index(q.z, p.sg_lo[2], p.sg_inv_h, p.sg_n[2]);
  return (uint32_t)(p.sg_base + ix + (int64_t)p.sg_n[0] * (iy + (int64_t)p.sg_n[1] * iz));
}

// The clump-cell grid of an instance, in registers: a caller loads it before the work that
// precedes the key (its loads then overlap that work).
struct ClumpGrid {
  double lo[3], inv_h, side_x, div;
  int32_t n[3];
  int64_t base;
};
__device__ __forceinline__ ClumpGrid clump_grid(const InstParams &p) {
  return ClumpGrid{{p.cg_lo[0], p.cg_lo[1], p.cg_lo[2]}, p.cg_inv_h, p.side_x, p.sphere_div,
                   {p.cg_n[0], p.cg_n[1], p.cg_n[2]}, p.cg_base};
}

// Clump-cell key of a clump centre for per-step binning by clump (kInvalidKey for a clump
// past side.x at t_new, which is not binned: cells.cu:65).
__device__ __forceinline__ uint32_t clump_key(const ClumpGrid &g, d3 rr, double t_new) {
  if (frozen(t_new, rr.x, g.side_x)) return kInvalidKey;  // not binned (cells.cu:65)
  const int ix = grid_index(rr.x, g.lo[0], g.inv_h, g.n[0]);
  const int iy = grid_index(rr.y, g.lo[1], g.inv_h, g.n[1]);
  const int iz = grid_index(rr.z, g.lo[2], g.inv_h, g.n[2]);
  return (uint32_t)(g.base + ix + (int64_t)g.n[0] * (iy + (int64_t)g.n[1] * iz));
}
__device__ __forceinline__ uint32_t clump_key(const InstParams &p, d3 rr, double t_new) {
  return clump_key(clump_grid(p), rr, t_new);
}

// Sphere centres rr + (rad/div) R (+-1,+-1,+-1), init_Verlet.cu:130-142.
__device__ __forceinline__ void sphere_centres(d3 rr, const double *R, double pos, d3 *s) {
  s[0] = make_d3(rr.x + pos * (R[0] + R[1] + R[2]), rr.y + pos * (R[3] + R[4] + R[5]),
                 rr.z + pos * (R[6] + R[7] + R[8]));
  s[1] = make_d3(rr.x + pos * (R[0] - R[1] - R[2]), rr.y + pos * (R[3] - R[4] - R[5]),
                 rr.z + pos * (R[6] - R[7] - R[8]));
  s[2] = make_d3(rr.x + pos * (-R[0] - R[1] + R[2]), rr.y + pos * (-R[3] - R[4] + R[5]),
                 rr.z + pos * (-R[6] - R[7] + R[8]));
  s[3] = make_d3(rr.x + pos * (-R[0] + R[1] - R[2]), rr.y + pos * (-R[3] + R[4] - R[5]),
                 rr.z + pos * (-R[6] + R[7] - R[8]));
}

// Centre of sphere i (0..3 = a..d) alone.  The sign factors are exact, and x + (-y) rounds
// as x - y, so this is bit-identical to the matching row of sphere_centres.
__device__ __forceinline__ d3 sphere_of(d3 rr, const double *R, double pos, int i) {
  const double sx = i < 2 ? 1.0 : -1.0;
  const double sy = (i == 0 || i == 3) ? 1.0 : -1.0;
  const double sz = (i == 0 || i == 2) ? 1.0 : -1.0;
  return make_d3(rr.x + pos * (sx * R[0] + sy * R[1] + sz * R[2]),
                 rr.y + pos * (sx * R[3] + sy * R[4] + sz * R[5]),
                 rr.z + pos * (sx * R[6] + sy * R[7] + sz * R[8]));
}

// The clump sweep's float record of a binned clump from its stage record g: its centre,
// relative to the instance's clump-grid origin, and radius (the record's: a binned clump is
// not frozen).
__device__ __forceinline__ void clump_fca(const double *g, const ClumpGrid &p,
                                          float4 *__restrict__ ca) {
  ca[0] = make_float4((float)(g[G_RR] - p.lo[0]), (float)(g[G_RR + 1] - p.lo[1]),
                      (float)(g[G_RR + 2] - p.lo[2]), (float)g[G_RAD]);
}

// The four sphere centres of a clump as floats relative to the clump-grid origin, from its
// stage record g (sphere_centres on the record's centre and rotation: the predict's FP64
// centres): sphere k at f[3k .. 3k + 2].
__device__ __forceinline__ void clump_fcs(const double *__restrict__ g, const ClumpGrid &p,
                                          float *f) {
  const d3 rr = make_d3(g[G_RR], g[G_RR + 1], g[G_RR + 2]);
  double R[9];
#pragma unroll
  for (int k = 0; k < 9; k++) R[k] = g[G_R + k];
  d3 s[4];
  sphere_centres(rr, R, g[G_RAD] / p.div, s);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    f[3 * k] = (float)(s[k].x - p.lo[0]);
    f[3 * k + 1] = (float)(s[k].y - p.lo[1]);
    f[3 * k + 2] = (float)(s[k].z - p.lo[2]);
  }
}

// Mass and moment of inertia (isotropic) of a clump of four spheres of radius rad and density
// rho, in the operation order of the library's setup (calculate_particle_parameters.cu:27-33;
// clump_mode 1: I = 4 m_s (2/5 r^2 + 2/3 d^2) for the sphere offset d = sqrt(3) r / div).
constexpr double kPiSetup = 3.1415926535897932;
__device__ __forceinline__ void clump_mass_inertia(int mode, double rad, double rho, double div,
                                                   double &mass, double &inertia) {
  const double aux1 = 4.0 * kPiSetup / 3.0, aux2 = 112.0 * kPiSetup / 15.0;
  const double rad3 = rad * rad * rad, rad5 = rad3 * rad * rad;
  mass = 4.0 * aux1 * rad3 * rho;
  if (mode == 0) {
    inertia = aux2 * rho * rad5;
  } else {
    const double ms = aux1 * rad3 * rho, d = sqrt(3.0) * rad / div;
    inertia = 4.0 * ms * (0.4 * rad * rad + (2.0 / 3.0) * d * d);
  }
}

// Centre of sphere i of the clump whose stage record is g (its centre and rotation rows);
// pos = the clump's radius / div (the record's radius: only records of clumps that are not
// frozen, whose radius is the one they had before freezing, are used this way).
__device__ __forceinline__ d3 rec_sphere(const double *__restrict__ g, double pos, int i) {
  double R[9];
#pragma unroll
  for (int k = 0; k < 9; k++) R[k] = g[G_R + k];
  return sphere_of(make_d3(g[G_RR], g[G_RR + 1], g[G_RR + 2]), R, pos, i);
}

// --------------------------------------------------------------------- k_predict ----

// Per clump, kept with the candidate lists: a bound on how far any of its spheres has moved
// since the list build, and the wall modules it can reach while the lists are valid.
struct ListTrack {
  double *dsp;
  uint8_t *wreach;
};

// init_Verlet.cu:46-114 on one clump in registers: position and velocity by the old
// accelerations, Rodrigues rotation of R by dOmega with column renormalisation, angular
// velocity.  Returns |dOmega|.
__device__ __forceinline__ double verlet_advance(d3 &rr, d3 &vv, d3 &ww, double *R, d3 aa, d3 ang,
                                                 double dt, double hdt) {
  // init_Verlet.cu:46-53
  rr.x += dt * (vv.x + hdt * aa.x);
  rr.y += dt * (vv.y + hdt * aa.y);
  rr.z += dt * (vv.z + hdt * aa.z);
  vv.x += dt * aa.x;
  vv.y += dt * aa.y;
  vv.z += dt * aa.z;
  // init_Verlet.cu:57-89: Rodrigues rotation by dOmega
  d3 dO = make_d3(dt * (ww.x + hdt * ang.x), dt * (ww.y + hdt * ang.y),
                  dt * (ww.z + hdt * ang.z));
  double mag = sqrt(dO.x * dO.x + dO.y * dO.y + dO.z * dO.z);
  double dR[9];
  if (mag > 0.0) {
    d3 nn = make_d3(dO.x / mag, dO.y / mag, dO.z / mag);
    double cdO = cos(mag), sdO = sin(mag), om = 1.0 - cdO;
    dR[0] = om * nn.x * nn.x + cdO;
    dR[1] = om * nn.x * nn.y - sdO * nn.z;
    dR[2] = om * nn.x * nn.z + sdO * nn.y;
    dR[3] = om * nn.y * nn.x + sdO * nn.z;
    dR[4] = om * nn.y * nn.y + cdO;
    dR[5] = om * nn.y * nn.z - sdO * nn.x;
    dR[6] = om * nn.z * nn.x - sdO * nn.y;
    dR[7] = om * nn.z * nn.y + sdO * nn.x;
    dR[8] = om * nn.z * nn.z + cdO;
  } else {
    dR[0] = 1.0; dR[1] = 0.0; dR[2] = 0.0;
    dR[3] = 0.0; dR[4] = 1.0; dR[5] = 0.0;
    dR[6] = 0.0; dR[7] = 0.0; dR[8] = 1.0;
  }
  // init_Verlet.cu:92-100: T = dR R (rows)
  double T[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      T[3 * i + j] = dR[3 * i] * R[j] + dR[3 * i + 1] * R[3 + j] + dR[3 * i + 2] * R[6 + j];
  // init_Verlet.cu:103-114: normalise each column
#pragma unroll
  for (int j = 0; j < 3; j++) {
    double norm = sqrt(T[j] * T[j] + T[3 + j] * T[3 + j] + T[6 + j] * T[6 + j]);
    R[j] = T[j] / norm;
    R[3 + j] = T[3 + j] / norm;
    R[6 + j] = T[6 + j] / norm;
  }
  ww.x += dt * ang.x;
  ww.y += dt * ang.y;
  ww.z += dt * ang.z;
  return mag;
}

// The 22-double stage record of a clump (what neighbours gather): centre, radius, velocity,
// angular velocity, rotation rows, acceleration.
__device__ __forceinline__ void fill_stage(double *__restrict__ g, d3 rr, double rad, d3 vv,
                                           d3 ww, const double *R, d3 aa) {
  g[G_RR + 0] = rr.x;
  g[G_RR + 1] = rr.y;
  g[G_RR + 2] = rr.z;
  g[G_RAD] = rad;
  g[G_VV + 0] = vv.x;
  g[G_VV + 1] = vv.y;
  g[G_VV + 2] = vv.z;
  g[G_WW + 0] = ww.x;
  g[G_WW + 1] = ww.y;
  g[G_WW + 2] = ww.z;
#pragma unroll
  for (int k = 0; k < 9; k++) g[G_R + k] = R[k];
  g[G_AA + 0] = aa.x;
  g[G_AA + 1] = aa.y;
  g[G_AA + 2] = aa.z;
}

// Candidate-list displacement bound of clump c since the list build, accumulated step by
// step: sum of |x' - x| + sqrt(3) pos |dOmega| (a sphere sits sqrt(3) pos from the centre,
// and a rotation by |dOmega| moves it at most that times as far), which is never below
// the distance any sphere has actually moved.  Frozen clumps do not move.  No-op without
// candidate lists (lt.dsp null).
__device__ __forceinline__ void track_update(ListTrack lt, int64_t c, d3 r0, d3 r1, double pos,
                                             double rot_step) {
  if (!lt.dsp) return;
  const double dx = r1.x - r0.x, dy = r1.y - r0.y, dz = r1.z - r0.z;
  const double d = sqrt(dx * dx + dy * dy + dz * dz) + 1.7320508075688772 * pos * rot_step;
  if (d > 0.0) lt.dsp[c] += d;
}

// Counting-sort counts of a clump's four sphere-cell keys, one atomic per binned sphere
// whose return value is the sphere's rank within its cell (k_scatter then places it at the
// cell's start + rank, with no atomics of its own), and the keys and ranks themselves.
// Measured against staging the counts of a workgroup in an LDS hash table first (one atomic
// per distinct cell): the staging cost more than the atomics it saved, in the force kernel
// (49.5 -> 63 us per launch at 250,000 clumps) and in the scatter (23.5 -> 42.6 us).
__device__ __forceinline__ void count_keys(const uint32_t *keys, uint32_t *__restrict__ counts,
                                           uint32_t *__restrict__ skey,
                                           uint32_t *__restrict__ srank, int64_t c) {
  uint32_t r[4];
#pragma unroll
  for (int s = 0; s < 4; s++) r[s] = keys[s] != kInvalidKey ? atomicAdd(&counts[keys[s]], 1u) : 0u;
  reinterpret_cast<uint4 *>(skey)[c] = make_uint4(keys[0], keys[1], keys[2], keys[3]);
  reinterpret_cast<uint4 *>(srank)[c] = make_uint4(r[0], r[1], r[2], r[3]);
}

// Sphere-cell keys of a clump's four spheres at the force-stage time t_new.  Sphere centres
// always use the radius the clump had before it froze (rad0): a frozen clump's centres are
// the stale ones of the step it crossed side.x (rr, R unchanged since).  Clumps past side.x
// at t_new are not binned (cells.cu:65): kInvalidKey.
__device__ __forceinline__ void clump_keys(const InstParams &p, d3 rr, const double *R,
                                           double rad0, double t_new, uint32_t *keys) {
  d3 sph[4];
  sphere_centres(rr, R, rad0 / p.sphere_div, sph);
  const bool out = frozen(t_new, rr.x, p.side_x);
#pragma unroll
  for (int s = 0; s < 4; s++) keys[s] = out ? kInvalidKey : sphere_key(p, sph[s]);
}

// mode 1: predict (init_Verlet) then stage and sphere-cell keys (into keys[4]); mode 3: the
// same without writing the predicted state back to st (per-step binning: a chunk's steps
// carry the state in the stage records, and its last step writes it back); mode 2: the
// same predict and stage without keys (candidate-list step), updating the displacement
// bound; mode 0: stage the current state only (initial force pass, avalanche_hst.cu:365-384,
// whose force stage sees the unchanged time), with keys.  t_old: the time at the start of the step (P.time, or
// within a chunk of list steps chunk_time(), which has the same sign: only comparisons of
// the time with 0 are made here).
__device__ __forceinline__ void predict_clump(double *__restrict__ st, double *__restrict__ sta,
                                              double *__restrict__ g, const InstParams &p,
                                              const int32_t *__restrict__ orig, uint32_t *keys,
                                              int64_t c, int64_t cap, int mode, ListTrack lt,
                                              double t_old, bool ckeys = false) {
  const double t_new = mode ? t_old + p.dt : t_old;

  d3 rr = make_d3(st[S_RX * cap + c], st[S_RY * cap + c], st[S_RZ * cap + c]);
  double R[9];
#pragma unroll
  for (int k = 0; k < 9; k++) R[k] = st[(S_R00 + k) * cap + c];
  double rad = sta[C_RAD * cap + c];
  const double rad0 = rad != 0.0 ? rad : sta[C_RAD0 * cap + c];  // radius before freezing
  d3 vv = make_d3(st[S_VX * cap + c], st[S_VY * cap + c], st[S_VZ * cap + c]);
  d3 ww = make_d3(st[S_WX * cap + c], st[S_WY * cap + c], st[S_WZ * cap + c]);
  double rot_step = 0.0;  // |dOmega| of this step
  const d3 r0 = rr;
  const d3 aa = make_d3(st[S_AX * cap + c], st[S_AY * cap + c], st[S_AZ * cap + c]);

  if (mode) {
    if (frozen(t_old, rr.x, p.side_x)) {
      rad = 0.0;
      sta[C_RAD * cap + c] = 0.0;
    } else {
      const d3 ang = make_d3(st[S_GX * cap + c], st[S_GY * cap + c], st[S_GZ * cap + c]);
      rot_step = verlet_advance(rr, vv, ww, R, aa, ang, p.dt, p.hdt);
      if (mode != 3) {
        st[S_RX * cap + c] = rr.x;
        st[S_RY * cap + c] = rr.y;
        st[S_RZ * cap + c] = rr.z;
        st[S_VX * cap + c] = vv.x;
        st[S_VY * cap + c] = vv.y;
        st[S_VZ * cap + c] = vv.z;
        st[S_WX * cap + c] = ww.x;
        st[S_WY * cap + c] = ww.y;
        st[S_WZ * cap + c] = ww.z;
#pragma unroll
        for (int k = 0; k < 9; k++) st[(S_R00 + k) * cap + c] = R[k];
      }
    }
  }

  fill_stage(g, rr, rad, vv, ww, R, aa);

  if (mode == 2) {
    track_update(lt, c, r0, rr, rad0 / p.sphere_div, rot_step);
    return;
  }

  if (ckeys)
    keys[0] = clump_key(p, rr, t_new);
  else
    clump_keys(p, rr, R, rad0, t_new, keys);
}

// A clump's state in registers: its stage record as read at the start of a list step
// (rr, vv, ww, R, aa = a_i; rad, rho, id), or its finished state handed from the force
// stage to the next step's predict inside a chunk of candidate-list steps (aa, ang = the
// new accelerations).
struct ClumpRegs {
  d3 rr, vv, ww, aa, ang;
  double R[9];
  double rad, rho;
  uint32_t id;
  bool frz;  // frozen at the force stage's time: no finish, no accelerations
};

// predict_clump mode 2 on a finished state in registers, writing only the stage record:
// inside a lean chunk a clump's centre, velocities, rotation and acceleration live in its
// record (the angular acceleration in st), and the chunk's last step writes them back to st.
__device__ __forceinline__ void predict_regs(ClumpRegs &s, double *__restrict__ sta,
                                             double *__restrict__ g, const InstParams &p,
                                             double rad, double rad0,
                                             int64_t c, int64_t cap, ListTrack lt,
                                             double t_old) {
  double rot_step = 0.0;
  const d3 r0 = s.rr;
  if (frozen(t_old, s.rr.x, p.side_x)) {
    rad = 0.0;
    sta[C_RAD * cap + c] = 0.0;
  } else {
    rot_step = verlet_advance(s.rr, s.vv, s.ww, s.R, s.aa, s.ang, p.dt, p.hdt);
  }
  fill_stage(g, s.rr, rad, s.vv, s.ww, s.R, s.aa);
  track_update(lt, c, r0, s.rr, rad0 / p.sphere_div, rot_step);
}

// The stage records of a workgroup's 256 clumps are assembled in LDS (rows padded to 23
// doubles against bank conflicts) and written as one contiguous span, instead of 22
// scattered 176-B-strided stores per lane.
constexpr int kPredictBlock = 256;
constexpr int kStagePad = kStageDoubles + 1;
__global__ __launch_bounds__(kPredictBlock) void k_predict(
    double *__restrict__ st, double *__restrict__ sta, double *__restrict__ stage,
    const int32_t *__restrict__ inst_of, const int32_t *__restrict__ orig,
    const InstParams *__restrict__ P, uint32_t *__restrict__ counts, uint32_t *__restrict__ skey,
    uint32_t *__restrict__ srank, int64_t n, int64_t cap, int mode, ListTrack lt, int ckeys,
    float4 *__restrict__ frec) {
  __shared__ double rec[kPredictBlock * kStagePad];
  const int64_t c0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kPredictBlock;
  const int64_t c = c0 + threadIdx.x;
  uint32_t keys[4] = {kInvalidKey, kInvalidKey, kInvalidKey, kInvalidKey};
  if (c < n) {
    const int inst = inst_of[c];
    while (true) {
      const int cur = __builtin_amdgcn_readfirstlane(inst);
      if (inst == cur) {
        predict_clump(st, sta, rec + threadIdx.x * kStagePad, P[cur], orig, keys, c, cap, mode,
                      lt, P[cur].time, ckeys != 0);
        break;
      }
    }
  }
  if (mode != 2 && c < n) {
    if (ckeys) {  // one clump key (clump-cell binning): skey / srank hold one word per clump
      skey[c] = keys[0];
      srank[c] = keys[0] != kInvalidKey ? atomicAdd(&counts[keys[0]], 1u) : 0u;
    } else {
      count_keys(keys, counts, skey, srank, c);
    }
  }
  __syncthreads();
  if (frec && c < n && keys[0] != kInvalidKey)  // the clump sweep's record (clump keys)
    clump_fca(rec + threadIdx.x * kStagePad, clump_grid(P[inst_of[c]]), frec + c);
  const int64_t rows = n - c0 < kPredictBlock ? n - c0 : kPredictBlock;
  double *dst = stage + c0 * kStageDoubles;
  for (int64_t k = threadIdx.x; k < rows * kStageDoubles; k += kPredictBlock)
    dst[k] = rec[(k / kStageDoubles) * kStagePad + k % kStageDoubles];
}

// ------------------------------------------------------------------------- scan -----
//
// Exclusive scan of the sphere-cell counts into starts, in two launches of a fixed grid of
// kScanGrid workgroups over the device-side cell count: k_count_sum sums each workgroup's
// share of the cells, k_count_scan scans the kScanGrid sums in every workgroup, then its
// share of the cells tile by tile (writing starts, zeroing the counts for the next binning).

constexpr int kScanGrid = 1024;
constexpr int kBsumStride = 32;  // words between the share sums the count pass adds into

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

// Exclusive block scan; *total receives the block sum.  lds needs kScanBlock/64 words.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *total, uint32_t *lds) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint32_t x = wave_incl_scan(v);
  __syncthreads();
  if (lane == 63) lds[wid] = x;
  __syncthreads();
  if (wid == 0) {
    uint32_t w = lane < nw ? lds[lane] : 0u;
    w = wave_incl_scan(w);
    if (lane < nw) lds[lane] = w;
  }
  __syncthreads();
  uint32_t off = wid ? lds[wid - 1] : 0u;
  *total = lds[nw - 1];
  return off + x - v;
}

// Cell range [lo, hi) of workgroup b: whole tiles of kScanTile cells, the same number of
// tiles for every workgroup.
__device__ __forceinline__ void scan_share(int64_t ncells, int64_t b, int64_t nb, int64_t &lo,
                                           int64_t &hi) {
  const int64_t tiles = (ncells + kScanTile - 1) / kScanTile;
  const int64_t per = (tiles + nb - 1) / nb;
  lo = b * per * kScanTile;
  hi = lo + per * kScanTile;
  if (lo > ncells) lo = ncells;
  if (hi > ncells) hi = ncells;
}

__global__ __launch_bounds__(kScanBlock) void k_count_sum(const uint32_t *__restrict__ counts,
                                                          uint32_t *__restrict__ bsum,
                                                          const int64_t *__restrict__ ncells_p) {
  __shared__ uint32_t lds[kScanBlock / 64];
  int64_t lo, hi;
  scan_share(*ncells_p, blockIdx.x, gridDim.x, lo, hi);
  uint32_t s = 0u;
  // lo is a multiple of kScanTile: 16-byte loads, coalesced across the workgroup
  for (int64_t i = lo + 4 * (int64_t)threadIdx.x; i < hi; i += 4 * kScanBlock) {
    if (i + 3 < hi) {
      const uint4 v = *reinterpret_cast<const uint4 *>(counts + i);
      s += v.x + v.y + v.z + v.w;
    } else {
      for (int64_t k = i; k < hi; k++) s += counts[k];
    }
  }
  uint32_t tot;
  block_excl_scan(s, &tot, lds);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// The next step's cell counts on per-step binning, in the sorted order of the step just
// binned: lane j takes the sphere at sorted position j, reads the key the force stage gave it
// for the next step and counts it with a returning atomic whose old value is the sphere's
// rank in its cell.  Spheres move little in a step, so the lanes of a wavefront count into a
// few adjacent counters and their atomics share cache lines (measured at 250,000 clumps:
// 9.2 us, against 17.0 us counting in clump order and about 16 us inside the force kernel).
// Every sphere with a key for the next step was binned in this one: a clump that is not
// binned is frozen (past side.x after t = 0), and a frozen clump neither moves nor thaws
// within a chunk (side.x only changes between chunks), so it stays unbinned.
//
// The same pass adds the counts of each share of cells of the scan that follows into
// bsum[share * kBsumStride] (the shares of k_count_sum, nb scan workgroups), so that scan
// needs no k_count_sum: in sorted order the lanes of a wavefront fall in one or two shares,
// and each wavefront adds once per share it meets.  The sums sit a cache line apart: atomics
// to one line are performed one after the other.  bsum must hold zeros (k_count_scan clears
// the buffer of the next step).
__global__ __launch_bounds__(256) void k_count_ranks(const uint32_t *__restrict__ skey,
                                                     const uint32_t *__restrict__ sid,
                                                     const uint32_t *__restrict__ starts,
                                                     const int64_t *__restrict__ ncells_p,
                                                     uint32_t *__restrict__ counts,
                                                     uint32_t *__restrict__ srank,
                                                     uint32_t *__restrict__ bsum, int32_t nb,
                                                     int64_t nitems) {
  const int64_t j = (int64_t)xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  const int64_t ncells = *ncells_p;
  bool todo = false;
  uint32_t key = kInvalidKey;
  // the binned items of this step (never more than nitems, the length of sid)
  const int64_t nbin = (int64_t)starts[ncells] < nitems ? (int64_t)starts[ncells] : nitems;
  if (j < nbin) {
    const uint32_t id = sid[j];
    key = skey[id];
    if (key != kInvalidKey) {
      srank[id] = atomicAdd(&counts[key], 1u);
      todo = true;
    }
  }
  // the scan workgroup whose share of cells holds key (scan_share's shares)
  const int64_t tiles = (ncells + kScanTile - 1) / kScanTile;
  const int64_t span = ((tiles + nb - 1) / nb) * kScanTile;
  const uint32_t w = todo ? (uint32_t)((int64_t)key / span) : 0u;
  const int lane = threadIdx.x & 63;
  while (true) {
    const unsigned long long b = __ballot(todo);
    if (!b) break;
    const int first = __ffsll((long long)b) - 1;
    const uint32_t wf = __shfl(w, first, 64);
    const bool mine = todo && w == wf;
    const unsigned long long mm = __ballot(mine);
    if (lane == first) atomicAdd(&bsum[(size_t)wf * kBsumStride], (uint32_t)__popcll(mm));
    todo = todo && !mine;
  }
}

// ----------------------------------------------------------------------- scatter ----

// One lane per sphere: the record (x, y, z, r) + id 4 m + s of each binned sphere at its
// cell's start + its rank within the cell (from the count's atomic).  The centre is the predict's (sphere_of on the stage record's centre and
// rotation is bit-identical to sphere_centres).  The sweep lists the touching spheres of
// binned spheres only; ccnt (if given) of a sphere that is not binned (its clump is past
// side.x) is set to 0 here, so every per-sphere count is written each binning.
constexpr int kScatterBlock = 256;
__global__ __launch_bounds__(kScatterBlock) void k_scatter(
    const uint32_t *__restrict__ skey, const uint32_t *__restrict__ srank,
    const uint32_t *__restrict__ starts, const double *__restrict__ stage,
    double4 *__restrict__ spos, uint32_t *__restrict__ sid,
    const int32_t *__restrict__ inst_of, const InstParams *__restrict__ P,
    int32_t *__restrict__ ccnt, DevDiag *__restrict__ diag, int64_t nsph) {
  const int64_t i = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kScatterBlock + threadIdx.x;
  if (i >= nsph) return;
  const uint32_t key = skey[i];
  if (key == kInvalidKey) {
    if (ccnt) ccnt[i] = 0;
    return;
  }
  const uint32_t dst = starts[key] + srank[i];
  if (dst >= (uint64_t)nsph) {  // counts and starts disagree: never expected; avf_step reports it
    atomicAdd(&diag->bin_broken, 1ull);
    return;
  }
  const double *g = stage + (i >> 2) * kStageDoubles;
  const double rad = g[G_RAD];  // binned: not frozen, the pre-freeze radius
  const d3 q = rec_sphere(g, rad / P[inst_of[i >> 2]].sphere_div, (int)(i & 3));
  spos[dst] = make_double4(q.x, q.y, q.z, rad);
  sid[dst] = (uint32_t)i;
}

// Up to kCopySets device-to-device copies of 4-byte words in one launch (the spatial
// reorder's buffer copies, the checkpoints of candidate-list chunks): 16-byte moves where
// both ends are 16-byte aligned, word moves for the rest.
constexpr int kCopySets = 4, kCopyBlock = 256, kCopyGrid = 2048;
struct CopySet {
  const uint32_t *src[kCopySets];
  uint32_t *dst[kCopySets];
  int64_t words[kCopySets];
};

__global__ __launch_bounds__(kCopyBlock) void k_copy_set(CopySet cs) {
  const int64_t stride = (int64_t)gridDim.x * kCopyBlock;
  const int64_t i0 = (int64_t)blockIdx.x * kCopyBlock + threadIdx.x;
#pragma unroll 1
  for (int k = 0; k < kCopySets; k++) {
    const uint32_t *src = cs.src[k];
    uint32_t *dst = cs.dst[k];
    const int64_t n = cs.words[k];
    int64_t done = 0;
    if (((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15u) == 0) {
      const int64_t n4 = n >> 2;
      for (int64_t i = i0; i < n4; i += stride)
        reinterpret_cast<uint4 *>(dst)[i] = reinterpret_cast<const uint4 *>(src)[i];
      done = n4 << 2;
    }
    for (int64_t i = done + i0; i < n; i += stride) dst[i] = src[i];
  }
}

// ------------------------------------------------------------------ boundaries -----

// force_hop.cu:9-244 for one nw (0: x walls, 1: y walls, 2: bottom gate while time < 0).
__device__ __forceinline__ void hop_walls(const InstParams &p, int nw, double time, d3 rr, d3 vv, d3 ww,
                          const double *R, double pos, double ri, double mass, double rb, d3 &ff, d3 &tq) {
  const double km = 0.5 * p.kappa_pw * mass, gm = 0.5 * p.gamma_pw * mass;
  const double xfront = p.hop_width, xrear = 0.0, yleft = p.yhop_l, yright = p.yhop_r;
  const double zbottom = p.hop_height;
  // every sphere centre lies within rb of rr: skip when no sphere can pass the tests below
  if (rr.z + rb < zbottom || rr.x - rb > p.hop_width) return;
  if (nw == 0 && rr.x - rb - xrear > ri && xfront - rr.x - rb > ri) return;
  if (nw == 1 && rr.y - rb - yleft > ri && yright - rr.y - rb > ri) return;
  if (nw == 2 && (!(time < 0.0) || rr.z - rb - zbottom > ri)) return;
#pragma unroll 1
  for (int ns = 0; ns < 4; ns++) {
    d3 r = sphere_of(rr, R, pos, ns), rrc = make_d3(0, 0, 0), nvec = make_d3(0, 0, 0);
    double dist, comp = 0.0;
    int touch = 0;
    if (r.z < zbottom || r.x > p.hop_width) continue;
    if (nw == 0) {
      dist = r.x - xrear;
      if (dist <= ri) {
        touch = 1;
        rrc = make_d3(xrear, r.y, r.z);
        nvec = make_d3(-1.0, 0.0, 0.0);
        comp = ri - dist;
      }
      dist = xfront - r.x;
      if (dist <= ri) {
        touch = 2;
        rrc = make_d3(xfront, r.y, r.z);
        nvec = make_d3(1.0, 0.0, 0.0);
        comp = ri - dist;
      }
    } else if (nw == 1) {
      dist = r.y - yleft;
      if (dist <= ri) {
        touch = 3;
        rrc = make_d3(r.x, yleft, r.z);
        nvec = make_d3(0.0, -1.0, 0.0);
        comp = ri - dist;
      }
      dist = yright - r.y;
      if (dist <= ri) {
        touch = 4;
        rrc = make_d3(r.x, yright, r.z);
        nvec = make_d3(0.0, 1.0, 0.0);
        comp = ri - dist;
      }
    } else {
      dist = r.z - zbottom;
      if (dist <= ri) {
        touch = (time < 0.0) ? 5 : 0;
        rrc = make_d3(r.x, r.y, zbottom);
        nvec = make_d3(0.0, 0.0, -1.0);
        comp = ri - dist;
      }
    }
    if (touch)
      wall_contact(rr, vv, ww, rrc, nvec, comp, km, gm, p.mu_pw, p.fr_pw, p.dt, mass, false, ff,
                   tq);
  }
}

// force_funnel.cu:9-213: converging side planes of the funnel (slope frame).
__device__ __forceinline__ void funnel_walls(const InstParams &p, d3 rr, d3 vv, d3 ww, const double *R, double pos, double ri,
                             double mass, double rb, d3 &ff, d3 &tq) {
  const double km = 0.5 * p.kappa_pw * mass, gm = 0.5 * p.gamma_pw * mass;
  const double sinfi = p.sinfi, cosfi = p.cosfi, tanfi = p.tanfi, cost = p.cost, sint = p.sint;
  {  // the plane distances below are 1-Lipschitz in the sphere centre
    const d3 qc = to_slope(rr, cost, sint);
    if (qc.x - rb > p.lfunnel - ri * sinfi) return;
    if (fabs(qc.y - qc.x * tanfi) * cosfi - rb > ri &&
        fabs(p.yfunnel - qc.y - qc.x * tanfi) * cosfi - rb > ri)
      return;
  }
#pragma unroll 1
  for (int ns = 0; ns < 4; ns++) {
    d3 q = to_slope(sphere_of(rr, R, pos, ns), cost, sint), rrc = make_d3(0, 0, 0), nvec = make_d3(0, 0, 0);
    double dist, comp = 0.0;
    int touch = 0;
    if (q.x > p.lfunnel - ri * sinfi) continue;
    dist = fabs(q.y - q.x * tanfi) * cosfi;
    if (dist <= ri) {
      touch = 1;
      d3 rc = make_d3(q.x + dist * sinfi, q.y - dist * cosfi, q.z);
      rrc = to_lab(rc, cost, sint);
      nvec = to_lab(make_d3(sinfi, -cosfi, 0.0), cost, sint);
      comp = ri - dist;
    }
    dist = fabs(p.yfunnel - q.y - q.x * tanfi) * cosfi;
    if (dist <= ri) {
      touch = 2;
      d3 rc = make_d3(q.x + dist * sinfi, q.y + dist * cosfi, q.z);
      rrc = to_lab(rc, cost, sint);
      nvec = to_lab(make_d3(sinfi, cosfi, 0.0), cost, sint);
      comp = ri - dist;
    }
    if (touch)
      wall_contact(rr, vv, ww, rrc, nvec, comp, km, gm, p.mu_pw, p.fr_pw, p.dt, mass, false, ff,
                   tq);
  }
}

// force_funnel.cu:216-404: vertical corner edges at the funnel exit.
__device__ __forceinline__ void funnel_corners(const InstParams &p, d3 rr, d3 vv, d3 ww, const double *R, double pos,
                               double ri, double mass, double rb, d3 &ff, d3 &tq) {
  const double km = 0.5 * p.kappa_pw * mass, gm = 0.5 * p.gamma_pw * mass;
  const double yfun_l = p.dy2, yfun_r = p.side_y + p.dy2, hsy = p.dy2 + p.side_y / 2.0;
  const double lf = p.lfunnel, cost = p.cost, sint = p.sint;
  {
    const d3 qc = to_slope(rr, cost, sint);
    if (qc.x + rb < lf - ri * p.sinfi || qc.x - rb > lf + ri) return;
  }
#pragma unroll 1
  for (int ns = 0; ns < 4; ns++) {
    d3 q = to_slope(sphere_of(rr, R, pos, ns), cost, sint);
    if (q.x < lf - ri * p.sinfi || q.x > lf + ri) continue;
    if (q.y > yfun_r || q.y < yfun_l) continue;
    double yc = (q.y <= hsy) ? yfun_l : yfun_r;
    double sep = ri, sep2 = sep * sep;
    d3 drr;
    drr.x = lf - q.x;
    double dist2 = drr.x * drr.x;
    if (dist2 > sep2) continue;
    drr.y = yc - q.y;
    dist2 += drr.y * drr.y;
    if (dist2 > sep2) continue;
    double dist = sqrt(dist2), comp = sep - dist;
    d3 nr = make_d3(drr.x / dist, drr.y / dist, 0.0);
    d3 rc = make_d3(q.x + ri * nr.x, q.y + ri * nr.y, q.z);
    d3 rrc = to_lab(rc, cost, sint), nvec = to_lab(nr, cost, sint);
    wall_contact(rr, vv, ww, rrc, nvec, comp, km, gm, p.mu_pw, p.fr_pw, p.dt, mass, true, ff, tq);
  }
}

// force_walls.cu:9-199: flume side walls beyond the funnel (caller skips frozen clumps).
__device__ __forceinline__ void side_walls(const InstParams &p, double time, d3 rr, d3 vv, d3 ww, const double *R, double pos,
                           double ri, double mass, double rb, d3 &ff, d3 &tq) {
  const double km = 0.5 * p.kappa_pw * mass;
  double gm = 0.5 * p.gamma_pw * mass, mu = p.mu_pw, cns_1 = p.fr_pw;
  const double yleft = p.dy2, yright = p.side_y + p.dy2;
  if (to_slope(rr, p.cost, p.sint).x + rb < p.lfunnel) return;
  if (rr.y - rb - yleft > ri && yright - rr.y - rb > ri) return;
#pragma unroll 1
  for (int ns = 0; ns < 4; ns++) {
    d3 r = sphere_of(rr, R, pos, ns), q = to_slope(r, p.cost, p.sint);
    d3 rrc = make_d3(0, 0, 0), nvec = make_d3(0, 0, 0);
    double dist, comp = 0.0;
    int touch = 0;
    if (q.x <= p.lfunnel) continue;
    dist = r.y - yleft;
    if (dist <= ri) {
      touch = 1;
      rrc = make_d3(r.x, yleft, r.z);
      nvec = make_d3(0.0, -1.0, 0.0);
      comp = ri - dist;
    }
    dist = yright - r.y;
    if (dist <= ri) {
      touch = 2;
      rrc = make_d3(r.x, yright, r.z);
      nvec = make_d3(0.0, 1.0, 0.0);
      comp = ri - dist;
    }
    if (!p.inter_flag && time > 0.0) gm = mu = cns_1 = 0.0;
    if (touch)
      wall_contact(rr, vv, ww, rrc, nvec, comp, km, gm, mu, cns_1, p.dt, mass, false, ff, tq);
  }
}

// force_flume.cu:10-182: inclined floor z = -x tan(theta).
__device__ __forceinline__ void flume_floor(const InstParams &p, d3 rr, d3 vv, d3 ww, const double *R, double pos, double ri,
                            double mass, double rb, d3 &ff, d3 &tq) {
  const double km = 0.5 * p.kappa_pw * mass, gm = 0.5 * p.gamma_pw * mass;
  const double sint = p.sint, cost = p.cost, tant = sint / cost;
  if (fabs(cost * (-tant * rr.x - rr.z)) - rb > ri) return;
#pragma unroll 1
  for (int ns = 0; ns < 4; ns++) {
    d3 r = sphere_of(rr, R, pos, ns);
    double rect = -tant * r.x - r.z, dist = fabs(cost * rect);
    if (!(dist <= ri)) continue;
    d3 nvec = make_d3(-sint, 0.0, -cost);
    d3 rrc = make_d3(r.x - dist * sint, r.y, r.z - dist * cost);
    wall_contact(rr, vv, ww, rrc, nvec, ri - dist, km, gm, p.mu_pw, p.fr_pw, p.dt, mass, false, ff,
                 tq);
  }
}

// force_dep.cu:10-178 with the deposit plane of avf_geometry.
__device__ __forceinline__ void deposit_floor(const InstParams &p, d3 rr, d3 vv, d3 ww, const double *R, double pos, double ri,
                              double mass, double rb, d3 &ff, d3 &tq) {
  const double km = 0.5 * p.kappa_pw * mass, gm = 0.5 * p.gamma_pw * mass;
  const double sind = p.dep_sin, cosd = p.dep_cos, tand = sind / cosd;
  if (cosd * (-(tand * (rr.x - p.dep_x0 - p.dep_x1) - (rr.z - p.dep_z0))) - rb > ri) return;
#pragma unroll 1
  for (int ns = 0; ns < 4; ns++) {
    d3 r = sphere_of(rr, R, pos, ns);
    double rect = tand * (r.x - p.dep_x0 - p.dep_x1) - (r.z - p.dep_z0), dist = cosd * (-rect);
    if (!(dist <= ri)) continue;
    d3 nvec = make_d3(sind, 0.0, -cosd);
    d3 rrc;
    rrc.x = r.x + dist * sind;
    rrc.y = r.y;
    rrc.z = tand * (rrc.x - p.dep_x0 - p.dep_x1) + p.dep_z0;
    wall_contact(rr, vv, ww, rrc, nvec, ri - dist, km, gm, p.mu_pw, p.fr_pw, p.dt, mass, false, ff,
                 tq);
  }
}

// Wall modules a clump centred at rr can still reach while none of its spheres moves more
// than `margin` (bit set = may touch): the early-out tests of the modules above, which are
// 1-Lipschitz in the centre, with the sphere offset bound rb widened by the margin.  The
// radius only ever drops (to 0 when the clump freezes), so each test takes whichever of ri
// and 0 makes it harder to pass (the funnel-mouth tests depend on the sign of sin(phi)).
// Time is not looked at (the hopper bottom counts as present).
enum WallReach : uint32_t {
  kReachHop = 1u, kReachFunnel = 2u, kReachCorners = 4u, kReachSide = 8u, kReachFlume = 16u,
  kReachDep = 32u, kReachAll = 63u
};
__device__ __forceinline__ uint32_t wall_reach(const InstParams &p, d3 rr, double ri,
                                               double rb) {
  uint32_t w = 0u;
  if (!(rr.z + rb < p.hop_height || rr.x - rb > p.hop_width) &&
      (!(rr.x - rb > ri && p.hop_width - rr.x - rb > ri) ||
       !(rr.y - rb - p.yhop_l > ri && p.yhop_r - rr.y - rb > ri) ||
       !(rr.z - rb - p.hop_height > ri)))
    w |= kReachHop;
  const d3 qc = to_slope(rr, p.cost, p.sint);
  if (!(qc.x - rb > p.lfunnel - (p.sinfi < 0.0 ? ri * p.sinfi : 0.0)) &&
      !(fabs(qc.y - qc.x * p.tanfi) * p.cosfi - rb > ri &&
        fabs(p.yfunnel - qc.y - qc.x * p.tanfi) * p.cosfi - rb > ri))
    w |= kReachFunnel;
  if (!(qc.x + rb < p.lfunnel - (p.sinfi > 0.0 ? ri * p.sinfi : 0.0) ||
        qc.x - rb > p.lfunnel + ri))
    w |= kReachCorners;
  if (!(qc.x + rb < p.lfunnel) &&
      !(rr.y - rb - p.dy2 > ri && p.side_y + p.dy2 - rr.y - rb > ri))
    w |= kReachSide;
  if (!(fabs(p.cost * (-(p.sint / p.cost) * rr.x - rr.z)) - rb > ri)) w |= kReachFlume;
  const double tand = p.dep_sin / p.dep_cos;
  if (!(p.dep_cos * (-(tand * (rr.x - p.dep_x0 - p.dep_x1) - (rr.z - p.dep_z0))) - rb > ri))
    w |= kReachDep;
  return w;
}

// ------------------------------------------------------------------- pp contacts ----

// Reference clump-cell of a clump centre by the cell_locate arithmetic (multiply by
// cs_inv, cells.cu:69-77) as the linear index cell_locate stores it (cells.cu:91; an index
// above ncell aliases into another cell exactly as in the reference).  Returns -1 if the
// reference would abort (negative index) or write outside its cell table.
__device__ __forceinline__ void ref_cell_raw(const InstParams &p, d3 r, int64_t &ii, int64_t &jj,
                                             int64_t &kk) {
  ii = (int64_t)(p.cs_inv[0] * (r.x - p.rr_min[0]));
  if (ii == p.ncell[0]) ii--;
  jj = (int64_t)(p.cs_inv[1] * (r.y - p.rr_min[1]));
  if (jj == p.ncell[1]) jj--;
  kk = (int64_t)(p.cs_inv[2] * (r.z - p.rr_min[2]));
  if (kk == p.ncell[2]) kk--;
}

__device__ __forceinline__ int64_t ref_cell_locate(const InstParams &p, d3 r) {
  int64_t ii, jj, kk;
  ref_cell_raw(p, r, ii, jj, kk);
  if (ii < 0 || jj < 0 || kk < 0) return -1;
  int64_t idx = ii + p.ncell[0] * (jj + kk * p.ncell[1]);
  if (idx >= p.ncell[0] * p.ncell[1] * p.ncell[2]) return -1;
  return idx;
}

// The cell of ref_cell_locate as (i, j, k) = (L mod nx, (L / nx) mod ny, L / (nx ny)); false
// where ref_cell_locate returns -1.  Inside the grid the raw indices are the answer, so the
// 64-bit divisions run only for the aliased indices of clumps beyond the grid.
__device__ __forceinline__ bool ref_cell_ijk(const InstParams &p, d3 r, int64_t &ai, int64_t &aj,
                                             int64_t &ak) {
  int64_t ii, jj, kk;
  ref_cell_raw(p, r, ii, jj, kk);
  if (ii < 0 || jj < 0 || kk < 0) return false;
  if (ii < p.ncell[0] && jj < p.ncell[1]) {
    if (kk >= p.ncell[2]) return false;  // L >= nx ny nz
    ai = ii;
    aj = jj;
    ak = kk;
    return true;
  }
  const int64_t nxy = p.ncell[0] * p.ncell[1];
  const int64_t L = ii + p.ncell[0] * (jj + kk * p.ncell[1]);
  if (L >= nxy * p.ncell[2]) return false;
  ak = L / nxy;
  const int64_t rem = L - ak * nxy;
  aj = rem / p.ncell[0];
  ai = rem - aj * p.ncell[0];
  return true;
}

constexpr int kSphList = 16;  // touching spheres of other clumps per sphere (k_sweep)

// Candidate list of one clump from the 4 sphere lists of k_sweep (skin > 0): every clump
// with a sphere pair within sep + skin, with the mask of those pairings, sorted by the
// reference key at the build (clump-cell offset, reference index; clumps the visit rule
// excludes at the build sort last, the force stage re-applies the rule every step).  Also
// resets the displacement tracking (centre at the build, rotation, bound).
constexpr int kVMax = 32;     // candidate clumps per clump
constexpr int kSphListV = 24; // candidate spheres per sphere
template <int BS>
__global__ __launch_bounds__(BS) void k_vmerge(
    const double *__restrict__ stage, const InstParams *__restrict__ P,
    const int32_t *__restrict__ inst_of, const int32_t *__restrict__ orig,
    const uint32_t *__restrict__ clist,
    const int32_t *__restrict__ ccnt, int32_t *__restrict__ vl_m, uint16_t *__restrict__ vl_mask,
    int32_t *__restrict__ vl_cnt, ListTrack lt, DevDiag *__restrict__ diag, int64_t n,
    int64_t cap, double skin) {
  __shared__ int32_t sm[kVMax][BS];
  __shared__ uint32_t sk[kVMax][BS];
  __shared__ uint16_t smask[kVMax][BS];
  const int lane = threadIdx.x;
  const int64_t c = (int64_t)xcd_block(blockIdx.x, gridDim.x) * BS + lane;
  if (c >= n) return;
  const InstParams &p = P[inst_of[c]];
  const double *gl = stage + c * kStageDoubles;
  const d3 rr = make_d3(gl[G_RR], gl[G_RR + 1], gl[G_RR + 2]);
  lt.dsp[c] = 0.0;
  {  // the wall modules within reach while no sphere moves skin/2 (frozen clumps: all)
    const double rad = gl[G_RAD];
    lt.wreach[c] = (uint8_t)(rad > 0.0 ? wall_reach(p, rr, rad,
                                                    1.7320508075688772 * (rad / p.sphere_div) *
                                                            (1.0 + 1e-12) +
                                                        1e-9 + skin)
                                       : kReachAll);
  }
  int64_t li = (int64_t)((rr.x - p.rr_min[0]) / p.cell_side[0]);
  if (li == p.ncell[0]) li--;
  int64_t lj = (int64_t)((rr.y - p.rr_min[1]) / p.cell_side[1]);
  if (lj == p.ncell[1]) lj--;
  int64_t lk = (int64_t)((rr.z - p.rr_min[2]) / p.cell_side[2]);
  if (lk == p.ncell[2]) lk--;
  int cnt = 0;
  bool overflow = false;
  for (int s = 0; s < 4; s++) {
    const int k = ccnt[4 * c + s];
    for (int e = 0; e < k; e++) {
      const uint32_t id = clist[(4 * c + s) * kSphListV + e];
      const int64_t cm = (int64_t)(id >> 2);
      const uint16_t bit = (uint16_t)(1u << (((id & 3u) << 2) | (uint32_t)s));
      int e2 = 0;
      for (; e2 < cnt; e2++)
        if (sm[e2][lane] == (int32_t)cm) break;
      if (e2 < cnt) {
        smask[e2][lane] = (uint16_t)(smask[e2][lane] | bit);
        continue;
      }
      const double *g = stage + cm * kStageDoubles;
      int64_t ai, aj, ak;
      uint32_t off = 27u;
      if (ref_cell_ijk(p, make_d3(g[G_RR], g[G_RR + 1], g[G_RR + 2]), ai, aj, ak)) {
        const int64_t di = ai - li, dj = aj - lj, dk = ak - lk;
        if (di >= -1 && di <= 1 && dj >= -1 && dj <= 1 && dk >= -1 && dk <= 1)
          off = (uint32_t)((di + 1) * 9 + (dj + 1) * 3 + (dk + 1));
      }
      const uint32_t key = (off << 27) | (uint32_t)orig[cm];
      if (cnt == kVMax) {
        overflow = true;
        continue;
      }
      int pos = cnt;
      while (pos > 0 && sk[pos - 1][lane] > key) {
        sk[pos][lane] = sk[pos - 1][lane];
        sm[pos][lane] = sm[pos - 1][lane];
        smask[pos][lane] = smask[pos - 1][lane];
        pos--;
      }
      sk[pos][lane] = key;
      sm[pos][lane] = (int32_t)cm;
      smask[pos][lane] = bit;
      cnt++;
    }
  }
  for (int e = 0; e < cnt; e++) {
    vl_m[e * cap + c] = sm[e][lane];
    vl_mask[e * cap + c] = smask[e][lane];
  }
  vl_cnt[c] = cnt;
  if (overflow) atomicAdd(&diag->neigh_overflow, 1ull);
}

// Per-lane columns only (no lane reads another lane's entries).  The pairing masks are
// built once the entries are sorted and reuse the keys' storage, which keeps a 64-clump
// workgroup at 20 KB of LDS (8 workgroups per CU, as many as the registers allow).
template <int MAXN, int BS>
struct PPShared {
  static_assert(MAXN >= 16, "the pairing masks reuse the key rows");
  int32_t m[MAXN][BS];         // storage index of each touching clump, sorted by key
  union {
    uint32_t key[MAXN][BS];    // (clump-cell offset << 27) | reference index
    uint32_t pm[16][BS];       // per pairing: bit e set when entry e touches in that pairing
  };
  uint16_t mask[MAXN][BS];     // bit (sdiv*4 + smod): sphere pairs in contact
};

// cell_dim.cu:55-60 margins and the regrid arithmetic of avalanche_hst.cu:311-333 when the
// regrid is due (or forced); then cc++ (avalanche_hst.cu:600, 852) when step is set.
__device__ __forceinline__ void finalize_params(InstParams &p, long long *a, int force, int step) {
  if (force || (p.regrid_interval > 0 && p.cc % p.regrid_interval == 0)) {
    const double rgyr = 0.5 * p.gyr_max, ddt = p.regrid_ddt_steps * p.dt;
    double rmin[3], rmax[3];
    for (int k = 0; k < 3; k++) {
      rmin[k] = order_val(a[k]);
      rmax[k] = order_val(a[3 + k]);
    }
    const double vxmax = order_val(a[6]), vzmin = order_val(a[7]);
    rmax[0] += 1.05 * rgyr + vxmax * ddt;
    rmin[0] -= 1.05 * rgyr;
    rmax[1] += 1.05 * rgyr;
    rmin[1] -= 1.05 * rgyr;
    rmax[2] += 1.05 * rgyr;
    rmin[2] -= 1.05 * rgyr - vzmin * ddt;
    const double dist = p.cell_dist;
    for (int k = 0; k < 3; k++) {
      p.rr_min[k] = rmin[k];
      p.rr_max[k] = rmax[k];
      p.length[k] = rmax[k] - rmin[k];
      p.ncell[k] = (int64_t)(p.length[k] / (dist * p.rmax));
      p.cell_side[k] = p.length[k] / ((double)p.ncell[k]);
      p.cs_inv[k] = 1.0 / p.cell_side[k];
    }
    for (int k = 0; k < 3; k++) {
      a[k] = 0x7fffffffffffffffll;
      a[3 + k] = (long long)0x8000000000000000ull;
    }
    a[6] = (long long)0x8000000000000000ull;
    a[7] = 0x7fffffffffffffffll;
  }
  if (step) p.cc += 1;
}

// Per wavefront: min / max of 8 order-preserving images over the lanes with active set,
// separately for each instance present (normally one), one atomic per instance and quantity.
// All 64 lanes of the wavefront must call it.  mode bit k: 1 = max, 0 = min.
__device__ __forceinline__ void wave_minmax8(long long *acc, int inst, const long long *vals,
                                             bool active, unsigned mode) {
  const int lane = threadIdx.x & 63;
  bool todo = active;
  while (true) {
    const unsigned long long b = __ballot(todo);
    if (!b) break;
    const int first = __ffsll((long long)b) - 1;
    const int cur = __shfl(inst, first);
    const bool mine = todo && inst == cur;
    for (int k = 0; k < 8; k++) {
      const bool mx = (mode >> k) & 1u;
      const long long idle = mx ? (long long)0x8000000000000000ull : 0x7fffffffffffffffll;
      long long x = mine ? vals[k] : idle;
      for (int off = 32; off > 0; off >>= 1) {
        const long long y = __shfl_xor(x, off);
        x = mx ? (x > y ? x : y) : (x < y ? x : y);
      }
      if (lane == first) {
        if (mx)
          atomicMax(&acc[8 * cur + k], x);
        else
          atomicMin(&acc[8 * cur + k], x);
      }
    }
    todo = todo && !mine;
  }
}

// The time of step si of a chunk whose P holds its first step, as far as the kernels use
// it: only its comparisons with 0 enter them (frozen(), the hopper gate, the side walls'
// friction switch).  The additions are k_finalize's, in its order, up to the point where the
// time is positive: adding dt > 0 keeps it positive, so the sign is that of the exact time.
__device__ __forceinline__ double chunk_time(const InstParams &q, int si) {
  double t = q.time;
  for (int k = 0; k < si && !(q.dt > 0.0 && t > 0.0); k++) t = t + q.dt;
  return t;
}

// Main per-clump force pass.  One lane per clump (storage order = spatial order).  The
// touching clumps come from k_sweep's per-sphere lists or, with VERLET, from the candidate
// list of k_vmerge, filtered every step by the same touching test and visit rule.
// time: the force stage's time (sys.time after the step's advance).
// State I/O of a clump (bit-identical results in every variant):
// * plain (not VERLET): own centre and velocities from st; the finish writes v, w, a, alpha.
// * VERLET, not NEXT (a rebuild step or the last step of a chunk): the own state (rr, v, w,
//   R, a_i, radius, density, index) from the stage record (own), which always holds it;
//   the finish writes rr, v, w, R, a, alpha to st (inside a LEAN chunk st keeps only alpha
//   current).
// * NEXT, not LEAN: as plain; the next step's predict then reads st and writes all of it.
// * NEXT and LEAN: own from the stage record; only alpha goes to st and the finished state
//   is left in out for the next step's predict (predict_regs), which writes the next
//   record: one read and one write of the state per clump-step.
// wreach (VERLET): the wall modules each clump can reach while the lists are valid.
template <int MAXN, int BS, bool VERLET, bool NEXT, bool LEAN, bool CLOSED, bool DEEP,
          bool CELLS>
__device__ __forceinline__ void force_clump(
    const InstParams &p, PPShared<MAXN, BS> &sh, int lane, int64_t c,
    double *__restrict__ st, const double *__restrict__ sta, const double *__restrict__ stage,
    const uint32_t *__restrict__ clist, const int32_t *__restrict__ ccnt,
    const int32_t *__restrict__ vl_m, const uint16_t *__restrict__ vl_mask,
    const int32_t *__restrict__ vl_cnt, const uint8_t *__restrict__ wreach,
    double *__restrict__ fout, DevDiag *__restrict__ diag, const int32_t *__restrict__ orig,
    int64_t cap, int initial, double time, const ClumpRegs &own, ClumpRegs &out) {
  static_assert(!NEXT || VERLET || CELLS, "the fused predict needs the record-resident state");
  static_assert(!CELLS || LEAN, "per-step binning keeps the state in the records");
  constexpr bool kRecState = VERLET || CELLS;  // the state lives in the stage records
  constexpr bool kOwnStage = kRecState && (LEAN || !NEXT);
  constexpr bool kRegsOut = NEXT && LEAN;
  constexpr bool kWriteRR = kRecState && !NEXT;
  const d3 rr = kOwnStage ? own.rr
                          : make_d3(st[S_RX * cap + c], st[S_RY * cap + c], st[S_RZ * cap + c]);
  d3 vv = kOwnStage ? own.vv : make_d3(st[S_VX * cap + c], st[S_VY * cap + c], st[S_VZ * cap + c]);
  d3 ww = kOwnStage ? own.ww : make_d3(st[S_WX * cap + c], st[S_WY * cap + c], st[S_WZ * cap + c]);
  double R[9];
#pragma unroll
  for (int k = 0; k < 9; k++) R[k] = kOwnStage ? own.R[k] : st[(S_R00 + k) * cap + c];
  const double rad = kOwnStage ? own.rad : sta[C_RAD * cap + c];
  // the radius before freezing: the radius itself until the clump freezes
  const double rad0 = rad != 0.0 ? rad : sta[C_RAD0 * cap + c];
  const double rho = kOwnStage ? own.rho : sta[C_RHO * cap + c];
  const double pos0 = rad0 / p.sphere_div;
  // bound on |sphere centre - rr| (sqrt(3) pos0; R has unit columns) with a margin far above
  // the rounding of the centres, for the boundary modules' early outs
  const double rb = 1.7320508075688772 * pos0 * (1.0 + 1e-12) + 1e-9;

  // the candidate list's length and first entry, loaded with the own record (the entry is
  // read whatever the length, so the loads do not wait for each other)
  const int kc0 = VERLET ? min(vl_cnt[c], kVMax) : 0;  // (the clump sweep counts past kVMax)
  const int32_t m00 = VERLET ? vl_m[c] : 0;
  const uint16_t b00 = VERLET ? vl_mask[c] : (uint16_t)0;
  const bool frz = frozen(time, rr.x, p.side_x);
  // the initial force pass (avalanche_hst.cu:365-384) runs force_part and then force_hop in
  // either phase, as the filling step does; the other boundary modules join on the steps
  const bool aval = p.phase && !initial;
  const int mods = (p.phase ? p.modules_aval : p.modules_fill) &
                   (initial ? (AVF_MOD_PP | AVF_MOD_HOP) : ~0);
  // on candidate-list steps, the wall modules this clump cannot reach before the lists are
  // rebuilt are skipped (each would return at its own early-out test)
  // (per-step binning by clump reads its lists the same way but has no reach bits)
  const uint32_t reach = (VERLET && !CELLS) ? (uint32_t)wreach[c] : (uint32_t)kReachAll;
  const double massw = kSphFact * rad * rad * rad * rho;  // force_hop.cu:45 order
  d3 ff = make_d3(0.0, 0.0, 0.0), tq = make_d3(0.0, 0.0, 0.0);

  if (aval) {  // avalanche_hst.cu:688-696
    if ((mods & AVF_MOD_HOP) && (reach & kReachHop))
#pragma unroll 1
      for (int nw = 0; nw < 3; nw++) hop_walls(p, nw, time, rr, vv, ww, R, pos0, rad, massw, rb, ff, tq);
    if ((mods & AVF_MOD_FUNNEL) && (reach & kReachFunnel))
      funnel_walls(p, rr, vv, ww, R, pos0, rad, massw, rb, ff, tq);
    if ((mods & AVF_MOD_FCORNERS) && (reach & kReachCorners))
      funnel_corners(p, rr, vv, ww, R, pos0, rad, massw, rb, ff, tq);
    if ((mods & AVF_MOD_WALLS) && !frz && (reach & kReachSide))
      side_walls(p, time, rr, vv, ww, R, pos0, rad, massw, rb, ff, tq);
  }

  if ((mods & AVF_MOD_PP) && !frz) {
    // ---- the touching clumps of the 4 sphere lists, sorted by the reference key -------
    int cnt = 0;
    bool overflow = false;
    // l's clump cell by DIVISION (force_part.cu:77-85)
    int64_t li = (int64_t)((rr.x - p.rr_min[0]) / p.cell_side[0]);
    if (li == p.ncell[0]) li--;
    int64_t lj = (int64_t)((rr.y - p.rr_min[1]) / p.cell_side[1]);
    if (lj == p.ncell[1]) lj--;
    int64_t lk = (int64_t)((rr.z - p.rr_min[2]) / p.cell_side[2]);
    if (lk == p.ncell[2]) lk--;
    {  // cell_locate's check of the clump itself: the reference exits (cells.cu:79-88)
      int64_t ii, jj, kk;
      ref_cell_raw(p, rr, ii, jj, kk);
      if (ii < 0 || jj < 0 || kk < 0) atomicAdd(&diag->cell_below, 1ull);
    }
    // one touching clump m (centre rm, reference index morig, touching pairings mask):
    // its entry under the reference's key, (clump-cell offset of m from l, m) ascending, by
    // insertion; a clump the reference's cell visit would not reach is counted and skipped
    auto place = [&](int64_t cm, d3 rm, uint32_t morig, uint32_t mask) {
      int64_t ai, aj, ak;
      if (!ref_cell_ijk(p, rm, ai, aj, ak)) {
        int64_t ii, jj, kk;
        ref_cell_raw(p, rm, ii, jj, kk);
        if (ii < 0 || jj < 0 || kk < 0)
          atomicAdd(&diag->cell_below, 1ull);  // cell_locate exits (cells.cu:79-88)
        else
          atomicAdd(&diag->cell_beyond, 1ull);  // past the reference's table
        return;
      }
      const int64_t di = ai - li, dj = aj - lj, dk = ak - lk;
      if (di < -1 || di > 1 || dj < -1 || dj > 1 || dk < -1 || dk > 1) {
        // outside the 27 clump cells the reference visits (force_part.cu:88-96)
        atomicAdd(&diag->cell_beyond, 1ull);
        return;
      }
      const uint32_t off = (uint32_t)((di + 1) * 9 + (dj + 1) * 3 + (dk + 1));
      const uint32_t key = (off << 27) | morig;
      if (cnt == MAXN) {
        overflow = true;
        return;
      }
      int pos = cnt;
      while (pos > 0 && sh.key[pos - 1][lane] > key) {
        sh.key[pos][lane] = sh.key[pos - 1][lane];
        sh.m[pos][lane] = sh.m[pos - 1][lane];
        sh.mask[pos][lane] = sh.mask[pos - 1][lane];
        pos--;
      }
      sh.key[pos][lane] = key;
      sh.m[pos][lane] = (int32_t)cm;
      sh.mask[pos][lane] = (uint16_t)mask;
      cnt++;
    };
    if constexpr (VERLET) {
      const int kc = kc0;
      // one candidate-list entry: the touching test of k_sweep on the same operands for each
      // candidate pairing, then its place
      auto visit = [&](int64_t cm, uint32_t bits, d3 rm, double radm, double idm,
                       const double *Rm) {
        if (frozen(time, rm.x, p.side_x)) return;  // not binned at this time (cells.cu:65)
        const double posm = radm / p.sphere_div;
        const double sep = rad + radm, sep2 = sep * sep;
        uint32_t mask = 0u;
        while (bits) {
          const int pr = __builtin_ctz(bits);
          bits &= bits - 1u;
          const int sdiv = pr >> 2, smod = pr & 3;
          const d3 q = sphere_of(rr, R, pos0, smod);
          const d3 o = sphere_of(rm, Rm, posm, sdiv);
          const double dx = o.x - q.x, dy = o.y - q.y, dz = o.z - q.z;
          if (dx * dx + dy * dy + dz * dz < sep2) mask |= 1u << pr;
        }
        if (mask) place(cm, rm, (uint32_t)__double_as_longlong(idm), mask);
      };
      // a candidate's centre, radius, id and rotation rows, loaded together
      auto gather = [&](int64_t cm, d3 &rm, double &radm, double &idm, double *Rm) {
        const double *g = stage + cm * kStageDoubles;
        rm = make_d3(g[G_RR], g[G_RR + 1], g[G_RR + 2]);
        radm = g[G_RAD];
        idm = __longlong_as_double((long long)(uint32_t)orig[cm]);
#pragma unroll
        for (int k = 0; k < 9; k++) Rm[k] = g[G_R + k];
      };
      int64_t cm_next = kc > 0 ? m00 : 0;
      uint32_t bits_next = kc > 0 ? b00 : 0u;
      if constexpr (!DEEP) {
        // the next entry is fetched one iteration ahead and a candidate's centre, radius, id
        // and sphere centres are loaded together, so an entry waits for one gather
        for (int e = 0; e < kc; e++) {
          const int64_t cm = cm_next;
          const uint32_t bits = bits_next;
          if (e + 1 < kc) {
            cm_next = vl_m[(e + 1) * cap + c];
            bits_next = vl_mask[(e + 1) * cap + c];
          }
          d3 rm;
          double radm, idm, Rm[9];
          gather(cm, rm, radm, idm, Rm);
          visit(cm, bits, rm, radm, idm, Rm);
        }
      } else {
        // small contexts (at most one wavefront per SIMD, registers to spare): the next
        // entry's record is gathered, and the one after's index fetched, while an entry is
        // tested
        d3 rm_n = make_d3(0.0, 0.0, 0.0);
        double radm_n = 0.0, idm_n = 0.0, Rm_n[9];
#pragma unroll
        for (int k = 0; k < 9; k++) Rm_n[k] = 0.0;
        if (kc > 0) gather(cm_next, rm_n, radm_n, idm_n, Rm_n);
        int64_t cm_nn = kc > 1 ? vl_m[cap + c] : 0;
        uint32_t bits_nn = kc > 1 ? vl_mask[cap + c] : 0u;
        for (int e = 0; e < kc; e++) {
          const int64_t cm = cm_next;
          const uint32_t bits = bits_next;
          const d3 rm = rm_n;
          const double radm = radm_n, idm = idm_n;
          double Rm[9];
#pragma unroll
          for (int k = 0; k < 9; k++) Rm[k] = Rm_n[k];
          if (e + 1 < kc) {
            cm_next = cm_nn;
            bits_next = bits_nn;
            gather(cm_next, rm_n, radm_n, idm_n, Rm_n);
            if (e + 2 < kc) {
              cm_nn = vl_m[(e + 2) * cap + c];
              bits_nn = vl_mask[(e + 2) * cap + c];
            }
          }
          visit(cm, bits, rm, radm, idm, Rm);
        }
      }
    } else {
      // the sphere lists of k_sweep: every entry is a touching pairing (l sphere s, m sphere
      // id & 3), added to m's entry or placed as a new one
      for (int s = 0; s < 4; s++) {
        const int k = ccnt[4 * c + s];
        for (int e = 0; e < k; e++) {
          const uint32_t id = clist[(4 * c + s) * kSphList + e];
          const int64_t cm = (int64_t)(id >> 2);
          const uint32_t bit = 1u << (((id & 3u) << 2) | (uint32_t)s);
          int e2 = 0;
          for (; e2 < cnt; e2++)
            if (sh.m[e2][lane] == (int32_t)cm) break;
          if (e2 < cnt) {
            sh.mask[e2][lane] = (uint16_t)(sh.mask[e2][lane] | bit);
            continue;
          }
          const double *g = stage + cm * kStageDoubles;
          place(cm, make_d3(g[G_RR], g[G_RR + 1], g[G_RR + 2]), (uint32_t)orig[cm], bit);
        }
      }
    }
    if (overflow) atomicAdd(&diag->neigh_overflow, 1ull);
    diag_max(&diag->max_neighbours, (unsigned long long)cnt);

    // ---- fold: pairings sph_id = 1..16 (m sphere outer, l sphere inner), then clump
    //      cells in (idel, jdel, kdel) order, then ascending m (force_part.cu:40-41,88-108)
    const double massl = kSphFact * rho * rad * rad * rad;
    const double kappa = p.kappa_pp, gamma = p.gamma_pp, mu = p.mu_pp, cns_1 = p.fr_pp;
    const double dt = p.dt;
    // transpose the entry masks into per-pairing entry sets (registers, then LDS) and note
    // which pairings have a contact at all
    uint32_t nz = 0u;
    {
      uint32_t pmr[16];
#pragma unroll
      for (int k = 0; k < 16; k++) pmr[k] = 0u;
      for (int e = 0; e < cnt; e++) {
        const uint32_t mk = sh.mask[e][lane];
        nz |= mk;
#pragma unroll
        for (int k = 0; k < 16; k++) pmr[k] |= ((mk >> k) & 1u) << e;
      }
#pragma unroll
      for (int k = 0; k < 16; k++) sh.pm[k][lane] = pmr[k];
    }
    // walk the set bits lane by lane so every lane evaluates only its own contacts
    int pr = 0;
    uint32_t cur = 0u;
    while (true) {
      if (cur == 0u) {
        if (nz == 0u) break;
        pr = __builtin_ctz(nz);
        nz &= nz - 1u;
        cur = sh.pm[pr][lane];
      }
      const int e = __builtin_ctz(cur);
      cur &= cur - 1u;
      const int sdiv = pr >> 2, smod = pr & 3;
      const d3 r1 = sphere_of(rr, R, pos0, smod);
      {
        const int64_t cm = sh.m[e][lane];
        const double *g = stage + cm * kStageDoubles;
        const d3 rrm = make_d3(g[G_RR], g[G_RR + 1], g[G_RR + 2]);
        const double rad_2 = g[G_RAD], rhom = sta[C_RHO * cap + cm];
        const d3 vvm = make_d3(g[G_VV], g[G_VV + 1], g[G_VV + 2]);
        const d3 wwm = make_d3(g[G_WW], g[G_WW + 1], g[G_WW + 2]);
        const d3 r2 = rec_sphere(g, rad_2 / p.sphere_div, sdiv);
        // force_part.cu:119-242
        const double massm = kSphFact * rhom * rad_2 * rad_2 * rad_2;
        const double meff = 0.5 * (massm + massl);
        const double km = 0.5 * kappa * meff, gm = 0.5 * gamma * meff;
        const double sep = rad + rad_2;
        d3 drr = make_d3(r2.x - r1.x, r2.y - r1.y, r2.z - r1.z);
        const double dist2 = drr.x * drr.x + drr.y * drr.y + drr.z * drr.z;
        const double dist = sqrt(dist2), comp = sep - dist;
        const d3 nvec = make_d3(drr.x / dist, drr.y / dist, drr.z / dist);
        const d3 rrc = make_d3(r1.x + rad * nvec.x, r1.y + rad * nvec.y, r1.z + rad * nvec.z);
        const d3 vl = point_vel(vv, ww, rrc, rr), vm = point_vel(vvm, wwm, rrc, rrm);
        const d3 dvvc = make_d3(vm.x - vl.x, vm.y - vl.y, vm.z - vl.z);
        Contact cl;
        if (contact_normal(cl, dvvc, nvec, comp, km, gm) > 0.0) {
          const d3 f = contact_friction(cl, mu, cns_1, dt, massl, true);
          accumulate(ff, tq, rrc, rr, f);
        }
      }
    }
  }

  if (!aval) {  // avalanche_hst.cu:460-461 (pp first, then hop, while filling)
    if ((mods & AVF_MOD_HOP) && (reach & kReachHop))
#pragma unroll 1
      for (int nw = 0; nw < 3; nw++) hop_walls(p, nw, time, rr, vv, ww, R, pos0, rad, massw, rb, ff, tq);
  } else {  // avalanche_hst.cu:704
    if ((mods & AVF_MOD_FLUME) && !frz && (reach & kReachFlume))
      flume_floor(p, rr, vv, ww, R, pos0, rad, massw, rb, ff, tq);
    if ((mods & AVF_MOD_DEP) && (reach & kReachDep))
      deposit_floor(p, rr, vv, ww, R, pos0, rad, massw, rb, ff, tq);
  }

  if (fout) {  // F / T of the step, kept for export (last step of a launch sequence)
    fout[F_FX * cap + c] = ff.x;
    fout[F_FY * cap + c] = ff.y;
    fout[F_FZ * cap + c] = ff.z;
    fout[F_TX * cap + c] = tq.x;
    fout[F_TY * cap + c] = tq.y;
    fout[F_TZ * cap + c] = tq.z;
  }

  if (frz) {
    if (!initial) vv = make_d3(0.0, 0.0, 0.0);  // finish_Verlet.cu:28-32
    if constexpr (kRegsOut) {
      out.rr = rr;
      out.vv = vv;
      out.ww = ww;
      out.aa = own.aa;  // no accelerations while frozen: a_i stays the last one computed
#pragma unroll
      for (int k = 0; k < 9; k++) out.R[k] = R[k];
      out.frz = true;
    } else {
      if constexpr (kWriteRR) {  // the chunk's last step: the record's state back to st
        st[S_RX * cap + c] = rr.x;
        st[S_RY * cap + c] = rr.y;
        st[S_RZ * cap + c] = rr.z;
        st[S_WX * cap + c] = ww.x;
        st[S_WY * cap + c] = ww.y;
        st[S_WZ * cap + c] = ww.z;
#pragma unroll
        for (int k = 0; k < 9; k++) st[(S_R00 + k) * cap + c] = R[k];
        st[S_AX * cap + c] = own.aa.x;
        st[S_AY * cap + c] = own.aa.y;
        st[S_AZ * cap + c] = own.aa.z;
      }
      if (!initial) {
        st[S_VX * cap + c] = 0.0;
        st[S_VY * cap + c] = 0.0;
        st[S_VZ * cap + c] = 0.0;
      }
    }
    return;
  }

  // get_accel.cu:27-71 (rad is the radius before freezing here)
  double mass;
  d3 I;
  if (CLOSED && p.mass_closed) {
    double ic;
    clump_mass_inertia(p.mass_closed - 1, rad, rho, p.sphere_div, mass, ic);
    I = make_d3(ic, ic, ic);
  } else {
    mass = sta[C_MASS * cap + c];
    I = make_d3(sta[C_IX * cap + c], sta[C_IY * cap + c], sta[C_IZ * cap + c]);
  }
  const double Cxy = (I.x - I.y) / I.z, Cyz = (I.y - I.z) / I.x, Czx = (I.z - I.x) / I.y;
  double adrag = p.drag * rad * rad / mass;
  const double vmag = sqrt(vv.x * vv.x + vv.y * vv.y + vv.z * vv.z);
  adrag *= vmag;
  d3 af;
  af.x = ff.x / mass - adrag * vv.x;
  af.y = ff.y / mass - adrag * vv.y;
  af.z = ff.z / mass - adrag * vv.z - p.grav;
  const d3 wb = make_d3(ww.x * R[0] + ww.y * R[3] + ww.z * R[6], ww.x * R[1] + ww.y * R[4] + ww.z * R[7],
                        ww.x * R[2] + ww.y * R[5] + ww.z * R[8]);
  const d3 tb = make_d3(tq.x * R[0] + tq.y * R[3] + tq.z * R[6], tq.x * R[1] + tq.y * R[4] + tq.z * R[7],
                        tq.x * R[2] + tq.y * R[5] + tq.z * R[8]);
  const d3 ab = make_d3(tb.x / I.x + Cyz * wb.y * wb.z, tb.y / I.y + Czx * wb.z * wb.x,
                        tb.z / I.z + Cxy * wb.x * wb.y);
  d3 gf;
  gf.x = ab.x * R[0] + ab.y * R[1] + ab.z * R[2];
  gf.y = ab.x * R[3] + ab.y * R[4] + ab.z * R[5];
  gf.z = ab.x * R[6] + ab.y * R[7] + ab.z * R[8];

  if (!initial) {  // finish_Verlet.cu:43-52
    const double hdt = p.hdt;
    const d3 ai = kOwnStage ? own.aa
                            : make_d3(st[S_AX * cap + c], st[S_AY * cap + c], st[S_AZ * cap + c]);
    const d3 gi = make_d3(st[S_GX * cap + c], st[S_GY * cap + c], st[S_GZ * cap + c]);
    vv.x += hdt * (af.x - ai.x);
    vv.y += hdt * (af.y - ai.y);
    vv.z += hdt * (af.z - ai.z);
    ww.x += hdt * (gf.x - gi.x);
    ww.y += hdt * (gf.y - gi.y);
    ww.z += hdt * (gf.z - gi.z);
    if constexpr (!kRegsOut) {
      if constexpr (kWriteRR) {
        st[S_RX * cap + c] = rr.x;
        st[S_RY * cap + c] = rr.y;
        st[S_RZ * cap + c] = rr.z;
#pragma unroll
        for (int k = 0; k < 9; k++) st[(S_R00 + k) * cap + c] = R[k];
      }
      st[S_VX * cap + c] = vv.x;
      st[S_VY * cap + c] = vv.y;
      st[S_VZ * cap + c] = vv.z;
      st[S_WX * cap + c] = ww.x;
      st[S_WY * cap + c] = ww.y;
      st[S_WZ * cap + c] = ww.z;
    }
  }
  if constexpr (kRegsOut) {
    out.rr = rr;
    out.vv = vv;
    out.ww = ww;
    out.aa = af;
    out.ang = gf;
#pragma unroll
    for (int k = 0; k < 9; k++) out.R[k] = R[k];
    out.frz = false;
  }
  // the accelerations become a_i of the next step (avalanche_hst.cu:844-850); inside a
  // lean chunk a_i travels in the stage record
  if constexpr (!kRegsOut) {
    st[S_AX * cap + c] = af.x;
    st[S_AY * cap + c] = af.y;
    st[S_AZ * cap + c] = af.z;
  }
  st[S_GX * cap + c] = gf.x;
  st[S_GY * cap + c] = gf.y;
  st[S_GZ * cap + c] = gf.z;
}

// Parameters of the clump's instance are wave-uniform in the waterfall below (one pass per
// distinct instance in the wavefront, normally one), so they are read with scalar loads
// instead of per-lane vector loads.
//
// NEXT (candidate-list steps inside a chunk): after finishing the step, the same lane runs
// the next step's predict on its clump (init_Verlet needs only the clump's own finished
// state) and writes the next stage record to stage_out, so a list step is one launch that
// reads and writes the carried state once.  Neighbours of this step read the records of
// `stage`, never stage_out.
//
// si: step si of a chunk of candidate-list steps whose P still holds the chunk's first
// step (k_finalize applies the chunk's per-step advances once, at its end): the step's
// time is chunk_time(P, si), its counter P.cc + si.
// time_ahead: the force stage runs at the step's time + dt (candidate-list steps advance
// the time at the end of the step instead of in k_scatter).  LEAN: see force_clump.  CLOSED:
// clump mass and inertia from radius and density for instances whose stored values are the
// closed forms (InstParams::mass_closed); a separate instantiation, since the code alone
// costs the other kernels registers.  DEEP: candidate records gathered one entry ahead, for
// contexts of at most one wavefront per SIMD (latency, not occupancy, bounds them).
template <int MAXN, int BS, bool VERLET, bool NEXT, bool LEAN, bool CLOSED, bool DEEP,
          bool CELLS>
__global__ __launch_bounds__(BS)
__attribute__((amdgpu_waves_per_eu(1, DEEP ? 1 : 8))) void k_force(
    double *__restrict__ st, double *__restrict__ sta, const double *__restrict__ stage,
    const int32_t *__restrict__ inst_of, const InstParams *__restrict__ P,
    const uint32_t *__restrict__ clist, const int32_t *__restrict__ ccnt,
    const int32_t *__restrict__ vl_m, const uint16_t *__restrict__ vl_mask,
    const int32_t *__restrict__ vl_cnt, double *__restrict__ fout, DevDiag *__restrict__ diag,
    long long *__restrict__ acc, double *__restrict__ stage_out,
    const int32_t *__restrict__ orig, ListTrack lt, int64_t n, int64_t cap, int initial,
    int time_ahead, int si, uint32_t *__restrict__ counts, uint32_t *__restrict__ skey,
    uint32_t *__restrict__ srank, float4 *__restrict__ frec, uint32_t *__restrict__ bsum,
    int32_t nb, const int64_t *__restrict__ ncells_p) {
  static_assert(!NEXT || VERLET || CELLS, "the fused predict needs the record-resident state");
  static_assert(sizeof(PPShared<MAXN, BS>) >= sizeof(double) * BS * kStagePad,
                "stage records of the workgroup are assembled in the pair scratch");
  __shared__ PPShared<MAXN, BS> sh;
  const int lane = threadIdx.x;
  const int64_t c0 = (int64_t)xcd_block_runs(blockIdx.x, gridDim.x, 16) * BS;
  const int64_t c = c0 + lane;
  const int inst = c < n ? inst_of[c] : -1;
  ClumpRegs cr, own;
  constexpr bool kOwnStage = (VERLET || CELLS) && (LEAN || !NEXT);
  constexpr bool kRegsOut = NEXT && LEAN;
  if (kOwnStage && c < n) {  // the own record, whole, in 16-byte loads
    const double2 *o = reinterpret_cast<const double2 *>(stage + c * kStageDoubles);
    const double2 a = o[0], b = o[1], v0 = o[2], v1 = o[3], w1 = o[4];
    const double2 r0 = o[5], r1 = o[6], r2 = o[7], r3 = o[8], r4 = o[9], r5 = o[10];
    own.rr = make_d3(a.x, a.y, b.x);
    own.rad = b.y;
    own.vv = make_d3(v0.x, v0.y, v1.x);
    own.ww = make_d3(v1.y, w1.x, w1.y);
    own.rho = sta[C_RHO * cap + c];
    own.id = (uint32_t)orig[c];
    own.R[0] = r0.x;
    own.R[1] = r0.y;
    own.R[2] = r1.x;
    own.R[3] = r1.y;
    own.R[4] = r2.x;
    own.R[5] = r2.y;
    own.R[6] = r3.x;
    own.R[7] = r3.y;
    own.R[8] = r4.x;
    own.aa = make_d3(r4.y, r5.x, r5.y);
  }
  if (c < n) {
    while (true) {
      const int cur = __builtin_amdgcn_readfirstlane(inst);
      if (inst == cur) {
        const InstParams &q = P[cur];
        const double t = chunk_time(q, si);
        force_clump<MAXN, BS, VERLET, NEXT, LEAN, CLOSED, DEEP, CELLS>(
            q, sh, lane, c, st, sta, stage, clist, ccnt, vl_m, vl_mask, vl_cnt, lt.wreach, fout,
            diag, orig, cap, initial, time_ahead ? t + q.dt : t, own, cr);
        break;
      }
    }
  }
  if (initial) return;  // the initial force pass (avalanche_hst.cu:365-384) is not a step
  // k_cellreduce's work on the steps the clump-cell regrid is due: the AABB of the new
  // centres and the velocity extremes (cell_dim.cu:27-52); k_finalize follows.
  bool due = false;
  long long v[8];
  if (c < n) {
    const InstParams &q = P[inst];
    due = q.regrid_interval > 0 && (q.cc + si) % q.regrid_interval == 0;
    if (due) {
      const d3 r = kRegsOut ? cr.rr
                            : make_d3(st[S_RX * cap + c], st[S_RY * cap + c], st[S_RZ * cap + c]);
      v[0] = order_key(r.x);
      v[1] = order_key(r.y);
      v[2] = order_key(r.z);
      v[3] = v[0];
      v[4] = v[1];
      v[5] = v[2];
      v[6] = order_key(kRegsOut ? cr.vv.x : st[S_VX * cap + c]);
      v[7] = order_key(kRegsOut ? cr.vv.z : st[S_VZ * cap + c]);
    }
  }
  if (__any(due)) wave_minmax8(acc, inst, v, due, 0x78u /* max: 3,4,5,6 */);
  if constexpr (NEXT) {
    __syncthreads();  // the pair scratch is free: it now holds the workgroup's records
    double *rec = reinterpret_cast<double *>(&sh);
    uint32_t keys[4] = {kInvalidKey, kInvalidKey, kInvalidKey, kInvalidKey};
    uint32_t rank = 0u;   // the clump's rank in its cell, stored last (the atomic's latency
    bool ranked = false;  // overlaps the record writes)
    ClumpGrid cg{};       // loaded before the predict, which then hides its latency
    if (c < n) {
      while (true) {
        const int cur = __builtin_amdgcn_readfirstlane(inst);
        if (inst == cur) {
          const InstParams &q = P[cur];
          // the next step starts at this step's force-stage time (P.time on per-step binning,
          // the chunk's step time + dt on candidate-list steps)
          const double t_old = CELLS ? q.time : chunk_time(q, si) + q.dt;
          cg = clump_grid(q);
          const double rad0 = own.rad != 0.0 ? own.rad : sta[C_RAD0 * cap + c];
          if constexpr (LEAN)
            predict_regs(cr, sta, rec + lane * kStagePad, q, own.rad, rad0, c,
                         cap, lt, t_old);
          else
            predict_clump(st, sta, rec + lane * kStagePad, q, orig, nullptr, c, cap, 2, lt,
                          t_old);
          if constexpr (CELLS && VERLET)
            keys[0] = clump_key(cg, cr.rr, t_old + q.dt);
          else if constexpr (CELLS)
            clump_keys(q, cr.rr, cr.R, rad0, t_old + q.dt, keys);
          break;
        }
      }
    }
    // the next step's keys (counted by k_count_ranks): the clump-cell key, or the four
    // sphere-cell keys
    if constexpr (CELLS && VERLET) {
      // the next step's clump key, its count and rank (the count's atomicAdd), and the next
      // scan's share sums (wave-aggregated, as k_count_ranks adds them)
      bool todo = false;
      if (c < n) {
        skey[c] = keys[0];
        if (keys[0] != kInvalidKey) {
          rank = atomicAdd(&counts[keys[0]], 1u);
          todo = true;
        }
      }
      ranked = todo;
      const int64_t tiles = (*ncells_p + kScanTile - 1) / kScanTile;
      const int64_t span = ((tiles + nb - 1) / nb) * kScanTile;
      const uint32_t w = todo ? (uint32_t)((int64_t)keys[0] / span) : 0u;
      while (true) {
        const unsigned long long b = __ballot(todo);
        if (!b) break;
        const int first = __ffsll((long long)b) - 1;
        const uint32_t wf = __shfl(w, first, 64);
        const bool mine = todo && w == wf;
        const unsigned long long mm = __ballot(mine);
        if (lane == first) atomicAdd(&bsum[(size_t)wf * kBsumStride], (uint32_t)__popcll(mm));
        todo = todo && !mine;
      }
    } else if constexpr (CELLS) {
      if (c < n)
        reinterpret_cast<uint4 *>(skey)[c] = make_uint4(keys[0], keys[1], keys[2], keys[3]);
    }
    __syncthreads();
    if constexpr (CELLS && VERLET)  // the clump sweep's float record of the next step
      if (c < n && keys[0] != kInvalidKey)
        clump_fca(rec + lane * kStagePad, cg, frec + c);
    const int64_t rows = n - c0 < BS ? n - c0 : BS;
    double *dst = stage_out + c0 * kStageDoubles;
    for (int64_t k = lane; k < rows * kStageDoubles; k += BS)
      dst[k] = rec[(k / kStageDoubles) * kStagePad + k % kStageDoubles];
    if (ranked) srank[c] = rank;
  }
}

// -------------------------------------------------------------- clump-cell regrid ---

// Block-level min / max of order-preserving images, then one atomic per block and quantity
// (lanes of other instances than the block's first contribute through their own atomics).
// vals: 8 images per lane; mode bit k: 1 = max, 0 = min.
__device__ __forceinline__ void block_minmax8(long long *acc_inst0, int inst0, int inst,
                                             long long *acc_lane, const long long *vals,
                                             bool active, unsigned mode) {
  __shared__ long long red[8][256];
  __shared__ int mixed;
  if (threadIdx.x == 0) mixed = 0;
  __syncthreads();
  if (active && inst != inst0) mixed = 1;
  for (int k = 0; k < 8; k++) {
    const bool mx = (mode >> k) & 1u;
    const long long idle = mx ? (long long)0x8000000000000000ull : 0x7fffffffffffffffll;
    red[k][threadIdx.x] = (active && inst == inst0) ? vals[k] : idle;
  }
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 8; k++) {
        const bool mx = (mode >> k) & 1u;
        const long long a = red[k][threadIdx.x], b = red[k][threadIdx.x + s];
        red[k][threadIdx.x] = mx ? (a > b ? a : b) : (a < b ? a : b);
      }
    __syncthreads();
  }
  if (threadIdx.x < 8) {
    const bool mx = (mode >> threadIdx.x) & 1u;
    if (mx)
      atomicMax(&acc_inst0[threadIdx.x], red[threadIdx.x][0]);
    else
      atomicMin(&acc_inst0[threadIdx.x], red[threadIdx.x][0]);
  }
  if (mixed && active && inst != inst0)
    for (int k = 0; k < 8; k++) {
      const bool mx = (mode >> k) & 1u;
      if (mx)
        atomicMax(&acc_lane[k], vals[k]);
      else
        atomicMin(&acc_lane[k], vals[k]);
    }
}

// cell_dim.cu:27-52 over every clump (frozen included) when cc % regrid_interval == 0
// (force = 1 runs it unconditionally).  acc per instance: min x,y,z, max x,y,z, max vx,
// min vz as order-preserving integers.  256 threads per block.
__global__ __launch_bounds__(256) void k_cellreduce(const double *__restrict__ st,
                                                    const int32_t *__restrict__ inst_of,
                                                    const InstParams *__restrict__ P,
                                                    long long *__restrict__ acc, int64_t n,
                                                    int64_t cap, int force, int32_t ninst) {
  // nothing to do unless some instance regrids this step (every regrid_interval steps)
  bool any = force != 0;
  for (int32_t i = 0; i < ninst && !any; i++)
    any = P[i].regrid_interval > 0 && P[i].cc % P[i].regrid_interval == 0;
  if (!any) return;
  const int64_t c0 = (int64_t)blockIdx.x * blockDim.x;
  const int64_t c = c0 + threadIdx.x;
  const int inst0 = inst_of[c0];
  const int inst = c < n ? inst_of[c] : inst0;
  const InstParams &p = P[inst];
  const bool due = force || (p.regrid_interval > 0 && p.cc % p.regrid_interval == 0);
  // whole-block early exit when the first instance is not due and the block is uniform
  const InstParams &p0 = P[inst0];
  const bool due0 = force || (p0.regrid_interval > 0 && p0.cc % p0.regrid_interval == 0);
  const int64_t clast = (c0 + blockDim.x - 1 < n) ? c0 + blockDim.x - 1 : n - 1;
  if (!due0 && inst_of[clast] == inst0) return;
  long long v[8];
  const bool active = c < n && due;
  if (active) {
    v[0] = order_key(st[S_RX * cap + c]);
    v[1] = order_key(st[S_RY * cap + c]);
    v[2] = order_key(st[S_RZ * cap + c]);
    v[3] = v[0];
    v[4] = v[1];
    v[5] = v[2];
    v[6] = order_key(st[S_VX * cap + c]);
    v[7] = order_key(st[S_VZ * cap + c]);
  }
  block_minmax8(acc + 8 * inst0, inst0, inst, acc + 8 * inst, v, active,
                0x78u /* max: 3,4,5,6 */);
}


// Per instance: finalize_params (step = 0 recomputes the grid without counting a step).
// advance_time: P.time += dt (candidate-list steps, whose force stage used P.time + dt).
// nsteps: the steps of a chunk of candidate-list steps, applied in order (a chunk ends on
// any step whose regrid is due, so only its last step can regrid).
__global__ void k_finalize(InstParams *__restrict__ P, long long *__restrict__ acc, int32_t ninst,
                           int force, int step, int advance_time = 0, int nsteps = 1) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ninst) return;
  InstParams &p = P[i];
  // counter, time and cc % regrid_interval (a running remainder: one 64-bit division per
  // launch) in registers; P is touched only on a regrid and at the end
  const int64_t ri = p.regrid_interval;
  int64_t cc = p.cc;
  int64_t r = ri > 0 ? ((cc % ri) + ri) % ri : 1;
  double t = p.time;
  const double dt = p.dt;
  for (int s = 0; s < nsteps; s++) {  // the steps of a chunk, in order
    if (force || (ri > 0 && r == 0)) {
      p.cc = cc;
      finalize_params(p, acc + 8 * i, 1, 0);
    }
    if (step) {
      cc += 1;
      if (ri > 0 && ++r == ri) r = 0;
    }
    if (advance_time) t = t + dt;
  }
  p.cc = cc;
  p.time = t;
}

// The parameter updates between two force stages of per-step binning, applied by
// k_count_scan: the previous step's clump-cell regrid when due and cc++ (k_finalize's work,
// avalanche_hst.cu:713-748, 852), then the time advance of the next step
// (avalanche_hst.cu:672, sys_par.time += sys_par.dt).
__device__ __forceinline__ void step_params(InstParams &p, long long *acc, int finalize_prev,
                                            int advance) {
  if (finalize_prev) {
    if (p.regrid_interval > 0 && p.cc % p.regrid_interval == 0) finalize_params(p, acc, 1, 0);
    p.cc += 1;
  }
  if (advance) p.time = p.time + p.dt;
}

// Exclusive scan of the counts into starts, the counts back to zero; starts[ncells] receives
// the total.  bsum: each workgroup's share sums, bstride words apart (k_count_sum: 1, or added
// up by the count pass: kBsumStride).
// Workgroup 0's first lane also applies step_params to every instance (no other kernel runs
// meanwhile; this kernel reads no parameter).  bsum_next (optional): the share sums the next
// count pass adds into (kBsumStride apart), zeroed here (each workgroup its own).
__global__ __launch_bounds__(kScanBlock) void k_count_scan(
    uint32_t *__restrict__ counts, const uint32_t *__restrict__ bsum,
    uint32_t *__restrict__ starts, const int64_t *__restrict__ ncells_p,
    InstParams *__restrict__ P, long long *__restrict__ acc, int32_t ninst, int finalize_prev,
    int advance, int bstride, uint32_t *__restrict__ bsum_next) {
  __shared__ uint32_t lds[kScanBlock / 64];
  if (bsum_next && threadIdx.x == 0) bsum_next[(size_t)blockIdx.x * kBsumStride] = 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0 && (finalize_prev || advance))
    for (int32_t i = 0; i < ninst; i++) step_params(P[i], acc + 8 * i, finalize_prev, advance);
  const int64_t ncells = *ncells_p;
  int64_t lo, hi;
  scan_share(ncells, blockIdx.x, gridDim.x, lo, hi);
  if (lo >= hi) {  // no cells in this share (the grid holds fewer than the launch covers)
    if (ncells == 0 && blockIdx.x == 0 && threadIdx.x == 0) starts[0] = 0u;
    return;
  }
  // tile by tile: wavefront w takes kScanTile / 4 cells as kRows rows of 64 (row k of lane l is
  // cell seg + 64 k + l: coalesced loads and stores), scans each row across the lanes, and the
  // rows and the 4 wavefronts are chained in order.  The first tile's counts are loaded
  // before the earlier shares' sums, so that both loads' latencies overlap.
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  constexpr int kRows = kScanTile / kScanBlock;
  uint32_t v[kRows];
  auto load_tile = [&](int64_t base) {
    const int64_t seg = base + (int64_t)wid * 64 * kRows;
#pragma unroll
    for (int k = 0; k < kRows; k++) {
      const int64_t i = seg + 64 * k + lane;
      v[k] = i < hi ? counts[i] : 0u;
    }
  };
  load_tile(lo);
  uint32_t s = 0u;  // sum of the earlier workgroups' counts
  for (int k = threadIdx.x; k < (int)blockIdx.x; k += kScanBlock) s += bsum[(size_t)k * bstride];
  uint32_t off;
  block_excl_scan(s, &off, lds);
  for (int64_t base = lo; base < hi; base += kScanTile) {
    const int64_t seg = base + (int64_t)wid * 64 * kRows;
    if (base != lo) load_tile(base);
    uint32_t inc[kRows], tot = 0u;
#pragma unroll
    for (int k = 0; k < kRows; k++) {
      inc[k] = wave_incl_scan(v[k]);
      tot += __shfl(inc[k], 63, 64);
    }
    __syncthreads();  // lds is free (the previous tile's reads are done)
    if (lane == 0) lds[wid] = tot;
    __syncthreads();
    uint32_t woff = off, tile = 0u;
    for (int w = 0; w < kScanBlock / 64; w++) {
      const uint32_t t = lds[w];
      if (w < wid) woff += t;
      tile += t;
    }
#pragma unroll
    for (int k = 0; k < kRows; k++) {
      const int64_t i = seg + 64 * k + lane;
      if (i < hi) {
        const uint32_t st0 = woff + inc[k] - v[k];
        starts[i] = st0;
        counts[i] = 0u;
      }
      woff += __shfl(inc[k], 63, 64);
    }
    off += tile;
  }
  if (threadIdx.x == 0 && hi == ncells) starts[ncells] = off;
}

// ------------------------------------------------------------------ contact sweep ---
//
// The touching spheres of every binned sphere, by bricks of sphere cells: a workgroup takes
// a brick of kSwBX x kSwBY x kSwBZ cells of an instance's grid and the one-cell halo around
// it, kSwRows rows of cells along x; a row is one contiguous range of the counting-sorted
// sphere array, so two cell starts per row bound it.  The workgroup copies the spheres of
// all rows to LDS as float records (centre relative to the halo origin, radius), lists the
// brick's own spheres, rebuilds each row's cell starts in LDS from the spheres' cells (the
// spheres of a row are sorted by cell), then each lane takes one own sphere and tests it
// against the spheres of the 27 cells around its own, reading LDS only.
//
// The listed set is exactly the set force_part.cu:140-157 evaluates: the pairs with
// ((dx^2 + dy^2) + dz^2) < sep^2 in FP64 on the sphere centres.  The float test decides the
// pairs whose distance is further than kSwMargin cell sides from sep (its rounding error is
// below a third of that margin while every coordinate lies within kSwFar cell sides of the
// halo origin); the pairs within the margin are decided by the FP64 test on the sorted
// centres themselves.  A halo holding more than kSwCap spheres, or a sphere further than
// kSwFar cell sides from the origin, is swept from global memory in FP64 (cell starts
// loaded, same test).  Each sphere lists the touching spheres of other clumps as
// (m << 2 | sm), at most CAP of them (the rest are counted in diag.neigh_overflow).  With
// skin > 0 (candidate lists) the test is d^2 < (sep + skin)^2 and the cells are at least
// 2 r_max + skin wide.  Brick b of all instances' bricks goes to logical workgroup b mod G
// (G = kSwGrid workgroups, more than stay resident, so that the hardware overlaps the
// phases of different bricks); runs of 8 consecutive bricks share an XCD and its L2 (the
// halos of neighbouring bricks overlap), successive runs go to successive XCDs.
constexpr int kSwBX = 16, kSwBY = 8, kSwBZ = 4;
constexpr int kSwHX = kSwBX + 2, kSwHY = kSwBY + 2, kSwHZ = kSwBZ + 2;
constexpr int kSwRows = kSwHY * kSwHZ;          // halo rows of a brick
constexpr int kSwOwnRows = kSwBY * kSwBZ;       // rows of the brick itself
constexpr int kSwRowStarts = kSwHX + 1;         // cell starts of a halo row and its end
constexpr int kSwBlock = 256;
constexpr int kSwCap = 704;                     // halo spheres staged in LDS
constexpr int kSwIt = 1;                        // halo loads in flight per thread
constexpr int kSwGrid = 8192;                   // workgroups (bricks in strides of kSwGrid)
constexpr float kSwMargin = 1.0f / 65536.0f;    // undecided band of the float test (cell sides)
constexpr double kSwFar = 20.0;                 // coordinate range of the float test (cell sides)
static_assert(kSwCap < 65536 && kSwRows < 256, "16-bit sphere and 8-bit row indices");
static_assert(kSwRows * kSwRowStarts * 4 <= kSwCap * 16, "global cell starts fit the records");

__device__ __forceinline__ int64_t bricks_of(const InstParams &p) {
  return (int64_t)((p.sg_n[0] + kSwBX - 1) / kSwBX) * ((p.sg_n[1] + kSwBY - 1) / kSwBY) *
         ((p.sg_n[2] + kSwBZ - 1) / kSwBZ);
}

// Index of the last entry of a[0..n) (ascending) that is <= v.
__device__ __forceinline__ int upper_row(const uint32_t *a, int n, uint32_t v) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a[mid] <= v)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

struct SweepLds {
  uint16_t hs[kSwRows][kSwRowStarts];  // first LDS sphere of each halo cell
  uint32_t lo[kSwRows];                // global index of each halo row's first sphere
  uint32_t hoff[kSwRows + 1];          // LDS index of each halo row's first sphere
  uint32_t ownp[kSwOwnRows + 1];       // global sweep: prefix of the brick rows' counts
  uint32_t nown;                       // own spheres listed in own[]
  uint32_t far;                        // a coordinate lies beyond the float test's range
  union {
    float4 pr[kSwCap];                 // centre - halo origin, radius + margin (float)
    uint32_t gs[kSwRows][kSwRowStarts];  // global sweep: first sorted sphere of each cell
  };
  uint32_t id[kSwCap];
  uint16_t own[kSwCap];                // LDS indices of the brick's own spheres
  uint8_t cx[kSwCap];                  // cell along x in the halo frame (0 .. kSwHX - 1)
  uint8_t row[kSwCap];                 // halo row of the sphere
};

// force_part.cu:140-157's test on the sorted centres (q: the listing sphere).
__device__ __forceinline__ bool touch_fp64(double4 q, double4 o, double skin) {
  const double sep = q.w + o.w + skin, sep2 = sep * sep;
  const double dx = o.x - q.x, dy = o.y - q.y, dz = o.z - q.z;
  return dx * dx + dy * dy + dz * dz < sep2;
}

// Halo row of the r9-th of the 27 cells' rows around halo row hr: hr + dy + kSwHY dz.
__device__ __forceinline__ int nb_row(int hr, int r9) {
  return hr + (r9 % 3 - 1) + kSwHY * (r9 / 3 - 1);
}

// Lists the touching spheres of own sphere t (LDS index) from the LDS records: cells
// lxc - 1 .. lxc + 1 of its 9 neighbour rows.  mh: the float test's margin (float, length).
// A pair within the margin is listed as kSwUndecided | its sorted index and decided by the
// FP64 test after the sweep (rare: the band is 2 mh wide), which keeps FP64 registers out
// of the candidate loop.
constexpr uint32_t kSwUndecided = 0x80000000u;  // sphere ids are below 2^31
template <int CAP>
__device__ __forceinline__ void sweep_lds(const SweepLds &sh, const double4 *__restrict__ spos,
                                          const uint32_t *__restrict__ sid, int t, double skin,
                                          float mh, uint32_t *__restrict__ clist,
                                          int32_t *__restrict__ ccnt, DevDiag *__restrict__ diag) {
  const float4 q = sh.pr[t];
  const uint32_t me = sh.id[t];
  const int hr = sh.row[t];
  const int lxc = sh.cx[t] < 1 ? 1 : (sh.cx[t] > kSwBX ? kSwBX : (int)sh.cx[t]);
  const float rqs = (q.w - mh) + (float)skin;  // own radius + skin; candidates hold r + mh
  uint32_t *out = clist + (int64_t)me * CAP;
  int cnt = 0;
  bool und = false;
  uint32_t na = sh.hs[nb_row(hr, 0)][lxc - 1], ne = sh.hs[nb_row(hr, 0)][lxc + 2];
#pragma unroll
  for (int r9 = 0; r9 < 9; r9++) {
    uint32_t j = na;
    const uint32_t e = ne;
    if (r9 < 8) {  // the next row's bounds load while this row is swept
      na = sh.hs[nb_row(hr, r9 + 1)][lxc - 1];
      ne = sh.hs[nb_row(hr, r9 + 1)][lxc + 2];
    }
    for (; j < e; j++) {
      const float4 o = sh.pr[j];
      const uint32_t id = sh.id[j];
      const float dx = o.x - q.x, dy = o.y - q.y, dz = o.z - q.z;
      const float d2 = __fmaf_rn(dx, dx, __fmaf_rn(dy, dy, dz * dz));
      const float hi = rqs + o.w;  // sep + margin
      if (d2 < hi * hi && (id ^ me) > 3u) {  // not surely apart, another clump
        const float lo = hi - 2.0f * mh;     // sep - margin
        uint32_t v = id;
        if (!(d2 < lo * lo)) {  // within the margin of sep
          const int h2 = nb_row(hr, r9);
          v = kSwUndecided | (sh.lo[h2] + j - sh.hoff[h2]);
          und = true;
        }
        if (cnt < CAP) out[cnt] = v;
        cnt++;
      }
    }
  }
  if (und) {  // the FP64 test for the undecided entries; the list is compacted in place
    const double4 q64 = spos[sh.lo[hr] + (uint32_t)t - sh.hoff[hr]];
    const int n = cnt < CAP ? cnt : CAP;
    int w = 0;
    unsigned long long decided = 0;
    for (int e = 0; e < n; e++) {
      uint32_t v = out[e];
      if (v & kSwUndecided) {
        decided++;
        const uint32_t jg = v & ~kSwUndecided;
        if (!touch_fp64(q64, spos[jg], skin)) continue;
        v = sid[jg];
      }
      out[w++] = v;
    }
    cnt -= n - w;
    atomicAdd(&diag->sweep_fp64, decided);
  }
  if (cnt > CAP) {
    atomicAdd(&diag->neigh_overflow, 1ull);
    cnt = CAP;
  }
  ccnt[me] = cnt;
}

// The same listing from global memory in FP64: sh.gs holds global indices into the sorted
// arrays; q, me: the listing sphere.
template <int CAP>
__device__ __forceinline__ void sweep_global(const SweepLds &sh, const double4 *__restrict__ spos,
                                             const uint32_t *__restrict__ sid, double4 q,
                                             uint32_t me, int hr, int lxc, double skin,
                                             uint32_t *__restrict__ clist,
                                             int32_t *__restrict__ ccnt,
                                             DevDiag *__restrict__ diag) {
  uint32_t *out = clist + (int64_t)me * CAP;
  int cnt = 0;
#pragma unroll 1
  for (int r9 = 0; r9 < 9; r9++) {
    const int h2 = nb_row(hr, r9);
    const uint32_t e = sh.gs[h2][lxc + 2];
    for (uint32_t j = sh.gs[h2][lxc - 1]; j < e; j++) {
      const double4 o = spos[j];
      const uint32_t id = sid[j];
      if (touch_fp64(q, o, skin) && (id ^ me) > 3u) {
        if (cnt < CAP) out[cnt] = id;
        cnt++;
      }
    }
  }
  if (cnt > CAP) {
    atomicAdd(&diag->neigh_overflow, 1ull);
    cnt = CAP;
  }
  ccnt[me] = cnt;
}

template <int CAP>
__global__ __launch_bounds__(kSwBlock) void k_sweep(
    const InstParams *__restrict__ P, int32_t ninst, const int64_t *__restrict__ nbricks_p,
    const uint32_t *__restrict__ starts, const double4 *__restrict__ spos,
    const uint32_t *__restrict__ sid, uint32_t *__restrict__ clist, int32_t *__restrict__ ccnt,
    DevDiag *__restrict__ diag, double skin) {
  __shared__ SweepLds sh;
  const int tid = threadIdx.x, lane = tid & 63;
  const int64_t nb = *nbricks_p;
  const int64_t G = gridDim.x, L = xcd_block_runs(blockIdx.x, gridDim.x, 8);
#pragma unroll 1
  for (int64_t b = L; b < nb; b += G) {
    __syncthreads();  // the previous brick's LDS is no longer read
    int inst = 0;
    if (ninst > 1) {  // last instance whose first brick is <= b
      int lo = 0, hi = ninst - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (P[mid].sg_bbase <= b)
          lo = mid;
        else
          hi = mid - 1;
      }
      inst = lo;
    }
    const InstParams &p = P[inst];
    const int nx = p.sg_n[0], ny = p.sg_n[1], nz = p.sg_n[2];
    const int64_t lb = b - p.sg_bbase;
    const int nbx = (nx + kSwBX - 1) / kSwBX, nby = (ny + kSwBY - 1) / kSwBY;
    const int x0 = (int)(lb % nbx) * kSwBX, y0 = (int)((lb / nbx) % nby) * kSwBY;
    const int z0 = (int)(lb / ((int64_t)nbx * nby)) * kSwBZ;
    const double lo_x = p.sg_lo[0], inv_h = p.sg_inv_h, h = p.sg_h;
    // the rows' bounds: cells x0 - 1 .. x0 + kSwBX of the row, clamped into [0, nx] (cells
    // outside the grid have no width; rows outside it are empty)
    uint32_t rlen = 0u;
    if (tid < kSwRows) {
      const int y = y0 - 1 + tid % kSwHY, z = z0 - 1 + tid / kSwHY;
      uint32_t a = 0u, e = 0u;
      if (y >= 0 && y < ny && z >= 0 && z < nz) {
        const int64_t row = p.sg_base + (int64_t)nx * (y + (int64_t)ny * z);
        const int xa = x0 - 1 < 0 ? 0 : x0 - 1;
        const int xe = x0 + kSwBX + 1 > nx ? nx : x0 + kSwBX + 1;
        a = starts[row + xa];
        e = starts[row + xe];
      }
      sh.lo[tid] = a;
      rlen = e - a;
    }
    if (tid < 64) {  // LDS offsets of the rows
      const uint32_t inc = wave_incl_scan(rlen);
      if (tid < kSwRows) sh.hoff[tid + 1] = inc;
      if (tid == 0) sh.hoff[0] = 0u;
    }
    if (tid == 0) {
      sh.nown = 0u;
      sh.far = 0u;
    }
    __syncthreads();
    const uint32_t H = sh.hoff[kSwRows];
    if (H == 0u) continue;
    const float mh = kSwMargin * (float)h;
    bool lds = H <= (uint32_t)kSwCap;
    if (lds) {
      // each row's thread marks its spheres with the row; empty cells start at the row end
      if (tid < kSwRows)
        for (uint32_t t = sh.hoff[tid]; t < sh.hoff[tid + 1]; t++) sh.row[t] = (uint8_t)tid;
      for (int t = tid; t < kSwRows * kSwRowStarts; t += kSwBlock)
        sh.hs[t / kSwRowStarts][t % kSwRowStarts] = (uint16_t)sh.hoff[t / kSwRowStarts + 1];
      __syncthreads();
      // the halo spheres: all of a thread's loads issued together (their latencies overlap),
      // then the float records, the cells along x and the list of the brick's own spheres
      const double ox = lo_x + (double)(x0 - 1) * h, oy = p.sg_lo[1] + (double)(y0 - 1) * h,
                   oz = p.sg_lo[2] + (double)(z0 - 1) * h, lim = kSwFar * h;
#pragma unroll 1
      for (uint32_t t0 = 0; t0 < H; t0 += kSwIt * kSwBlock) {
        int hrs[kSwIt];
        double4 ov[kSwIt];
        uint32_t iv[kSwIt];
#pragma unroll
        for (int q = 0; q < kSwIt; q++) {
          const uint32_t t = t0 + tid + q * kSwBlock;
          hrs[q] = 0;
          if (t < H) {
            hrs[q] = sh.row[t];
            const uint32_t j = sh.lo[hrs[q]] + (t - sh.hoff[hrs[q]]);
            ov[q] = spos[j];
            iv[q] = sid[j];
          }
        }
#pragma unroll
        for (int q = 0; q < kSwIt; q++) {
          const uint32_t t = t0 + tid + q * kSwBlock;
          bool own = false;
          if (t < H) {
            const double ux = ov[q].x - ox, uy = ov[q].y - oy, uz = ov[q].z - oz;
            if (!(fabs(ux) <= lim && fabs(uy) <= lim && fabs(uz) <= lim)) sh.far = 1u;
            sh.pr[t] = make_float4((float)ux, (float)uy, (float)uz, (float)ov[q].w + mh);
            sh.id[t] = iv[q];
            int lx = grid_index(ov[q].x, lo_x, inv_h, nx) - (x0 - 1);
            lx = lx < 0 ? 0 : (lx > kSwHX - 1 ? kSwHX - 1 : lx);
            sh.cx[t] = (uint8_t)lx;
            const int hy = hrs[q] % kSwHY, hz = hrs[q] / kSwHY;
            own = hy >= 1 && hy <= kSwBY && hz >= 1 && hz <= kSwBZ && lx >= 1 && lx <= kSwBX;
          }
          const unsigned long long m = __ballot(own);
          if (m) {  // the wave's own spheres appended together
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            uint32_t base = 0u;
            if (lane == 0) base = atomicAdd(&sh.nown, (uint32_t)__popcll(m));
            base = __shfl(base, 0, 64);
            if (own) sh.own[base + rank] = (uint16_t)t;
          }
        }
      }
      __syncthreads();
      lds = sh.far == 0u;
      if (lds) {
        // cell starts from the sorted cells: sphere t starts every cell after its
        // predecessor's
        for (uint32_t t = tid; t < H; t += kSwBlock) {
          const int hr = sh.row[t];
          const int prev = t == sh.hoff[hr] ? -1 : (int)sh.cx[t - 1];
          for (int k = prev + 1; k <= (int)sh.cx[t]; k++) sh.hs[hr][k] = (uint16_t)t;
        }
      }
    }
    if (!lds) {
      if (tid == 0) atomicAdd(&diag->sweep_global, 1ull);
      for (int t = tid; t < kSwRows * kSwRowStarts; t += kSwBlock) {
        const int hr = t / kSwRowStarts, k = t % kSwRowStarts;
        const int y = y0 - 1 + hr % kSwHY, z = z0 - 1 + hr / kSwHY;
        uint32_t v = 0u;
        if (y >= 0 && y < ny && z >= 0 && z < nz) {
          int x = x0 - 1 + k;
          x = x < 0 ? 0 : (x > nx ? nx : x);
          v = starts[p.sg_base + (int64_t)nx * (y + (int64_t)ny * z) + x];
        }
        sh.gs[hr][k] = v;
      }
      __syncthreads();
      if (tid < 64) {  // spheres of the brick rows (cells 1 .. kSwBX of the halo frame)
        uint32_t len = 0u;
        if (tid < kSwOwnRows) {
          const int hr = (1 + tid % kSwBY) + kSwHY * (1 + tid / kSwBY);
          len = sh.gs[hr][kSwBX + 1] - sh.gs[hr][1];
        }
        const uint32_t inc = wave_incl_scan(len);
        if (tid < kSwOwnRows) sh.ownp[tid + 1] = inc;
        if (tid == 0) sh.ownp[0] = 0u;
      }
    }
    __syncthreads();
    if (lds) {
      const uint32_t K = sh.nown;
      for (uint32_t k = tid; k < K; k += kSwBlock)
        sweep_lds<CAP>(sh, spos, sid, sh.own[k], skin, mh, clist, ccnt, diag);
    } else {
      const uint32_t K = sh.ownp[kSwOwnRows];
      for (uint32_t k = tid; k < K; k += kSwBlock) {
        const int o = upper_row(sh.ownp, kSwOwnRows, k);
        const int hr = (1 + o % kSwBY) + kSwHY * (1 + o / kSwBY);
        const uint32_t j = sh.gs[hr][1] + (k - sh.ownp[o]);
        const double4 v = spos[j];
        int lxc = grid_index(v.x, lo_x, inv_h, nx) - x0 + 1;
        lxc = lxc < 1 ? 1 : (lxc > kSwBX ? kSwBX : lxc);
        sweep_global<CAP>(sh, spos, sid, v, sid[j], hr, lxc, skin, clist, ccnt, diag);
      }
    }
  }
}

// ----------------------------------------------------------------- clump sweep -------
//
// Per-step binning by clump: one key per clump (the clump-cell grid cell of its centre, cell
// side >= the largest clump's bounding diameter, so the clumps of any touching pair sit in
// neighbouring cells), counted and scanned like the sphere keys (a quarter of the keys, a
// fifth of the cells).  The predict writes each binned clump's float record in storage order
// (its centre relative to the grid origin and its radius, as floats: 16 B, read by every
// bounding-sphere test); k_cscatter moves it into cell order, with the clump's storage index
// and key beside it (one 8-B pair).  The sphere centres, needed only for the few candidates
// that pass that test, are computed then from the clump's stage record, as floats relative
// to the grid origin.  k_csweep then lists, for every binned clump, the clumps of its 27
// neighbour cells that may touch it, with the pairings that may: a float test with a margin
// far above
// its rounding error (centre distance against the bounding radii, then the 16 sphere pairs),
// so the listed pairings are a superset of the touching ones.  The force kernel re-tests
// every listed pairing with force_part.cu:140-157's FP64 test on the stage records' centres
// (as on candidate-list steps), so the contacts, their order and every bit of the forces are
// those of the sphere-level binning.

constexpr int kCsBlock = 256;
constexpr int kCq = 8;  // candidates queued per lane between the sweep's two passes

// One lane per clump (storage order): the centre and radius of its float record, its storage
// index and its key to its place in cell order (its cell's start + its rank); every clump's
// list is emptied (the sweep appends to it).  A clump that is not binned (past side.x)
// keeps an empty list.
__global__ __launch_bounds__(256) void k_cscatter(const uint32_t *__restrict__ ckey,
                                                  const uint32_t *__restrict__ crank,
                                                  const uint32_t *__restrict__ starts,
                                                  const float4 *__restrict__ frec,
                                                  float4 *__restrict__ sca,
                                                  uint2 *__restrict__ cik,
                                                  int32_t *__restrict__ vl_cnt,
                                                  DevDiag *__restrict__ diag, int64_t n) {
  const int64_t c = (int64_t)xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (c >= n) return;
  const uint32_t key = ckey[c], rank = crank[c];
  const float4 v = frec[c];
  vl_cnt[c] = 0;
  if (key == kInvalidKey) return;
  const uint32_t dst = starts[key] + rank;
  if (dst < (uint64_t)n) {
    sca[dst] = v;
    cik[dst] = make_uint2((uint32_t)c, key);
  } else {  // counts and starts disagree: never expected; avf_step reports it
    atomicAdd(&diag->bin_broken, 1ull);
  }
}

// Candidate pairings of two clumps whose bounding spheres overlap: bit (sm * 4 + sl) for
// sphere sm of o and sl of q when their centres are closer than lim (r_q + r_o + margin).
__device__ __forceinline__ uint32_t cand_mask(const float *qs, const float *o, float lim) {
  const float lim2 = lim * lim;
  uint32_t mask = 0u;
#pragma unroll
  for (int sm = 0; sm < 4; sm++)
#pragma unroll
    for (int sl = 0; sl < 4; sl++) {
      const float ex = o[3 * sm] - qs[3 * sl], ey = o[3 * sm + 1] - qs[3 * sl + 1],
                  ez = o[3 * sm + 2] - qs[3 * sl + 2];
      if (__fmaf_rn(ex, ex, __fmaf_rn(ey, ey, ez * ez)) < lim2) mask |= 1u << (sm * 4 + sl);
    }
  return mask;
}

// The float test's margin for the own clump (xq: a bound on its coordinates' magnitude) and
// a candidate with first record quarter oa: above the test's rounding error for the two
// records, which scales with the largest coordinate magnitude.  Every sphere centre of a
// clump lies within r (1 + sqrt(3) / div) of its centre, so |centre| + r * bfac bounds the
// magnitude of every coordinate of its record.
__device__ __forceinline__ float pair_margin(float mh0, float xq, float4 oa, float bfac) {
  const float eo = fmaxf(fmaxf(fabsf(oa.x), fabsf(oa.y)), fabsf(oa.z)) + oa.w * bfac;
  return fmaxf(mh0, (xq + eo) * 0x1p-21f);
}

// The candidate lists of every binned clump, one lane per clump in cell order.  Every pair of
// neighbouring clumps is tested once, by the lane of its clump earlier in cell order: its
// partner lies in the same row after it, in the row above or in the three rows of the next
// layer, so a lane reads its clump's record, storage index and key (consecutive lanes:
// consecutive addresses), the first and last clump of the three cells x - 1 .. x + 1 of
// those 5 rows (each a contiguous range of the cell order), then walks the 5 ranges reading
// the records from global memory (lanes of a wavefront sit in the same or neighbouring
// cells, so they read the same ranges).  Two passes, so that the lanes of a wavefront test
// sphere pairs together instead of each whenever one of them meets a candidate: the
// bounding-sphere test queues each candidate that passes (kCq per lane in LDS; a full queue
// is emptied on the spot), then the 16 sphere pairs of the queued candidates.  A pair that
// may touch is entered in both clumps' lists (vl_m / vl_mask, entry-major as the candidate
// lists of k_vmerge; slots taken by atomicAdd on the list lengths, which the scatter
// emptied), the partner's entry with the transposed pairing mask: both tests are symmetric
// (float subtraction is, and the margin is built from both records alike), so a clump's
// list holds the same pairings as a test from its own side.  At most kVMax entries per
// clump are kept (the rest are counted in diag.neigh_overflow; the length may count past
// kVMax).
__global__ __launch_bounds__(kCsBlock) void k_csweep(
    const InstParams *__restrict__ P, int32_t ninst, const int64_t *__restrict__ ncells_p,
    const uint32_t *__restrict__ starts, const float4 *__restrict__ sca,
    const double *__restrict__ stage, const uint2 *__restrict__ cik,
    int32_t *__restrict__ vl_m, uint16_t *__restrict__ vl_mask, int32_t *__restrict__ vl_cnt,
    DevDiag *__restrict__ diag, int64_t cap, int64_t n) {
  __shared__ uint32_t q[kCq][kCsBlock];
  __shared__ uint32_t rb[9][kCsBlock], re[9][kCsBlock];
  const int tid = threadIdx.x;
  const int64_t k = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kCsBlock + tid;
  if (k >= n) return;
  // the lane's clump is loaded together with the binned count instead of after it, so that
  // the two latencies overlap.  A lane past the binned clumps (its entry left by an earlier
  // step, or never written) walks no candidate: any key gives rows inside the instance's
  // grid (the inside test below), and its ranges are emptied by the binned count.
  const uint2 ik = cik[k];
  const float4 qa = sca[k];
  const uint32_t sk = starts[*ncells_p];
  const int64_t nbin = sk < (uint64_t)n ? (int64_t)sk : n;  // clumps binned this step
  const uint32_t lim = k < nbin ? (uint32_t)nbin : 0u;
  const uint32_t key = ik.y, me = ik.x;
  int inst = 0;
  if (ninst > 1) {  // last instance whose first cell is <= key
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (P[mid].cg_base <= (int64_t)key)
        lo = mid;
      else
        hi = mid - 1;
    }
    inst = lo;
  }
  while (true) {  // instance parameters wave-uniform (one pass per instance in the wavefront)
    const int cur = __builtin_amdgcn_readfirstlane(inst);
    if (inst != cur) continue;
    const InstParams &p = P[cur];
    const int nx = p.cg_n[0], ny = p.cg_n[1], nz = p.cg_n[2];
    const int64_t local = (int64_t)key - p.cg_base, yz = local / nx;
    const bool inside = local >= 0 && local < (int64_t)nx * ny * nz;
    const int x = (int)(local - yz * nx), y = (int)(yz % ny), z = (int)(yz / ny);
    const int xl = x > 0 ? x - 1 : 0, xh = x + 2 < nx ? x + 2 : nx;
    // the 9 ranges (cells outside the grid are empty), all loaded together
    uint32_t b9[9], e9[9];
#pragma unroll
    for (int r9 = 0; r9 < 9; r9++) {
      const int yy = y + r9 % 3 - 1, zz = z + r9 / 3 - 1;
      b9[r9] = e9[r9] = 0u;
      if (r9 >= 4 && inside && yy >= 0 && yy < ny && zz >= 0 && zz < nz) {
        const int64_t row = p.cg_base + (int64_t)nx * (yy + (int64_t)ny * zz);
        b9[r9] = starts[row + xl];
        e9[r9] = starts[row + xh];
      }
    }
#pragma unroll
    for (int r9 = 0; r9 < 9; r9++) {  // within the sorted clumps whatever happens
      rb[r9][tid] = b9[r9] < lim ? b9[r9] : lim;
      re[r9][tid] = e9[r9] < lim ? e9[r9] : lim;
    }
    const float mh0 = (float)p.cg_mh, bfac = (float)p.cg_bfac;
    // a bound on the own record's coordinates (as pair_margin bounds the candidate's)
    const float xq = fmaxf(fmaxf(fabsf(qa.x), fabsf(qa.y)), fabsf(qa.z)) + qa.w * bfac;
    // the own sphere centres, from the own stage record when the first candidate is queued
    // (most clumps of a loose packing have none)
    const ClumpGrid cgp = clump_grid(p);
    float qs[12];
    bool own = false;
    int nq = 0;
    auto drain = [&]() {
      if (nq > 0 && !own) {
        clump_fcs(stage + (int64_t)me * kStageDoubles, cgp, qs);
        own = true;
      }
      for (int t = 0; t < nq; t++) {
        const uint32_t j = q[t][tid], m = cik[j].x;
        const float4 oa = sca[j];
        float os[12];
        clump_fcs(stage + (int64_t)m * kStageDoubles, cgp, os);
        const uint32_t mask =
            cand_mask(qs, os, qa.w + oa.w + pair_margin(mh0, xq, oa, bfac));
        if (!mask) continue;
        // the pair's entry in both lists: the own one with this mask, the candidate's with
        // its transpose (its spheres first); slots from the lists' lengths
        uint32_t mt = 0u;
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
          for (int b = 0; b < 4; b++) mt |= ((mask >> (a * 4 + b)) & 1u) << (b * 4 + a);
        const int s0 = atomicAdd(&vl_cnt[me], 1), s1 = atomicAdd(&vl_cnt[m], 1);
        if (s0 < kVMax) {
          vl_m[s0 * cap + me] = (int32_t)m;
          vl_mask[s0 * cap + me] = (uint16_t)mask;
        } else {
          atomicAdd(&diag->neigh_overflow, 1ull);
        }
        if (s1 < kVMax) {
          vl_m[s1 * cap + m] = (int32_t)me;
          vl_mask[s1 * cap + m] = (uint16_t)mt;
        } else {
          atomicAdd(&diag->neigh_overflow, 1ull);
        }
      }
      nq = 0;
    };
    // bounding-sphere test of candidate j (first record quarter oa); queues it when it passes
    auto test = [&](uint32_t j, float4 oa) {
      const float dx = oa.x - qa.x, dy = oa.y - qa.y, dz = oa.z - qa.z;
      const float bb = __fmaf_rn(qa.w + oa.w, bfac, pair_margin(mh0, xq, oa, bfac));
      if (j == (uint32_t)k || !(__fmaf_rn(dx, dx, __fmaf_rn(dy, dy, dz * dz)) < bb * bb)) return;
      if (nq == kCq) drain();
      q[nq++][tid] = j;
    };
    // each pair is found once, from its clump earlier in cell order: the candidates after k,
    // which lie in the own row (from k + 1), the row above and the three rows of the next
    // layer (rows 4 .. 8)
#pragma unroll 1
    for (int r9 = 4; r9 < 9; r9++) {
      uint32_t j = rb[r9][tid];
      if (r9 == 4 && j <= (uint32_t)k) j = (uint32_t)k + 1u;
      const uint32_t e = re[r9][tid];
      for (; j + 1 < e; j += 2) {  // two candidates' records in flight
        const float4 oa = sca[j], ob = sca[j + 1];
        test(j, oa);
        test(j + 1, ob);
      }
      if (j < e) test(j, sca[j]);
    }
    drain();
    break;
  }
}

// --------------------------------------------------------------- sphere-cell grid ---

// AABB of clump centres and max |v| per instance (acc: 8 per instance as in
// k_cellreduce; slot 6 holds max |v|^2 as an ordered double).  256 threads per block.
// kSgItems clumps per thread (strided by the block size): a block's clumps belong to its
// first clump's instance but where a block straddles instances, and those of other
// instances go straight to their own atomics, so a launch makes a few hundred atomics on
// each instance's accumulators instead of one per block of 256 clumps (atomics on one cache
// line are performed one after the other).
constexpr int kSgItems = 8;
__global__ __launch_bounds__(256) void k_sgrid_reduce(const double *__restrict__ st,
                                                      const int32_t *__restrict__ inst_of,
                                                      long long *__restrict__ acc, int64_t n,
                                                      int64_t cap) {
  const int64_t c0 = (int64_t)blockIdx.x * blockDim.x * kSgItems;
  const int inst0 = inst_of[c0];
  long long v[8] = {0x7fffffffffffffffll, 0x7fffffffffffffffll, 0x7fffffffffffffffll,
                    (long long)0x8000000000000000ull, (long long)0x8000000000000000ull,
                    (long long)0x8000000000000000ull, (long long)0x8000000000000000ull,
                    0x7fffffffffffffffll};
  bool active = false;
  for (int k = 0; k < kSgItems; k++) {
    const int64_t c = c0 + threadIdx.x + (int64_t)k * blockDim.x;
    if (c >= n) break;
    const double x = st[S_RX * cap + c], y = st[S_RY * cap + c], z = st[S_RZ * cap + c];
    const double vx = st[S_VX * cap + c], vy = st[S_VY * cap + c], vz = st[S_VZ * cap + c];
    const long long kx = order_key(x), ky = order_key(y), kz = order_key(z);
    const long long kv = order_key(vx * vx + vy * vy + vz * vz);
    const int inst = inst_of[c];
    if (inst == inst0) {
      v[0] = kx < v[0] ? kx : v[0];
      v[1] = ky < v[1] ? ky : v[1];
      v[2] = kz < v[2] ? kz : v[2];
      v[3] = kx > v[3] ? kx : v[3];
      v[4] = ky > v[4] ? ky : v[4];
      v[5] = kz > v[5] ? kz : v[5];
      v[6] = kv > v[6] ? kv : v[6];
      active = true;
    } else {
      long long *a = acc + 8 * inst;
      atomicMin(&a[0], kx);
      atomicMin(&a[1], ky);
      atomicMin(&a[2], kz);
      atomicMax(&a[3], kx);
      atomicMax(&a[4], ky);
      atomicMax(&a[5], kz);
      atomicMax(&a[6], kv);
    }
  }
  block_minmax8(acc + 8 * inst0, inst0, inst0, acc + 8 * inst0, v, active, 0x78u);
}

// Lays out every instance's sphere grid (cell side h >= 2 * largest sphere radius, so any
// touching pair sits in neighbouring cells; indices are clamped into the grid, which keeps
// that property for spheres outside it) and assigns key ranges.  Coarsens h until the
// total fits the cell capacity.  Single thread.
__global__ void k_sgrid_finalize(InstParams *__restrict__ P, long long *__restrict__ acc,
                                 int32_t ninst, int64_t cell_cap, double factor, double horizon,
                                 int64_t *__restrict__ ncells_out, double skin,
                                 int64_t *__restrict__ nbricks_out, int64_t *__restrict__ cg_out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double scale = 1.0;
  bool fit = false;
  for (int attempt = 0; attempt < 64 && !fit; attempt++) {
    int64_t total = 0;
    for (int i = 0; i < ninst; i++) {
      InstParams &p = P[i];
      long long *a = acc + 8 * i;
      const double vmax = sqrt(order_val(a[6]));
      const double margin = p.sg_reach + vmax * horizon * p.dt + 1e-9;
      // side >= 2 r_max (touching pairs in the 27 cells around a sphere), >= 2 r_max + skin
      // with candidate lists
      const double h = fmax(2.0 * p.sg_rsph * factor, 2.0 * p.sg_rsph + skin) * scale * (1.0 + 1e-9);
      int64_t nd[3];
      for (int k = 0; k < 3; k++) {
        const double lo = order_val(a[k]) - margin, hi = order_val(a[3 + k]) + margin;
        double len = hi - lo;
        if (!(len > 0.0)) len = h;
        int64_t m = (int64_t)ceil(len / h);
        if (m < 1) m = 1;
        if (m > (1 << 20)) m = 1 << 20;
        nd[k] = m;
        p.sg_lo[k] = lo;
      }
      p.sg_h = h;
      p.sg_inv_h = 1.0 / h;
      p.sg_n[0] = (int32_t)nd[0];
      p.sg_n[1] = (int32_t)nd[1];
      p.sg_n[2] = (int32_t)nd[2];
      p.sg_base = total;
      total += nd[0] * nd[1] * nd[2];
    }
    if (total <= cell_cap) {
      *ncells_out = total;
      fit = true;
    } else {
      scale *= cbrt((double)total / (double)cell_cap) * 1.05;
    }
  }
  if (!fit) {  // one cell per instance always fits (ninst <= clumps <= cell_cap)
    for (int i = 0; i < ninst; i++) {
      P[i].sg_n[0] = P[i].sg_n[1] = P[i].sg_n[2] = 1;
      P[i].sg_base = i;
    }
    *ncells_out = ninst;
  }
  int64_t bricks = 0;  // the contact sweep's bricks, instance after instance
  for (int i = 0; i < ninst; i++) {
    P[i].sg_bbase = bricks;
    bricks += bricks_of(P[i]);
  }
  *nbricks_out = bricks;
  if (cg_out) {
    // the clump-cell grid over the same box: cell side >= twice the largest bounding radius
    // (sg_reach), so the clumps of a touching pair sit in neighbouring cells
    scale = 1.0;
    fit = false;
    for (int attempt = 0; attempt < 64 && !fit; attempt++) {
      int64_t total = 0;
      for (int i = 0; i < ninst; i++) {
        InstParams &p = P[i];
        const double h = fmax(2.0 * p.sg_reach, p.sg_h) * scale * (1.0 + 1e-9);
        double ext = h;
        for (int k = 0; k < 3; k++) {
          const double len = (double)p.sg_n[k] * p.sg_h;
          int64_t m = (int64_t)ceil(len / h);
          if (m < 1) m = 1;
          if (m > (1 << 20)) m = 1 << 20;
          p.cg_n[k] = (int32_t)m;
          p.cg_lo[k] = p.sg_lo[k];
          ext = fmax(ext, (double)m * h);
        }
        p.cg_h = h;
        p.cg_inv_h = 1.0 / h;
        // the float test's margin: far above its rounding error for coordinates within the
        // grid (the sweep widens it for coordinates beyond, from the step's largest one)
        p.cg_mh = fmax(h * 0x1p-12, (ext + h) * 0x1p-20);
        p.cg_bfac = (1.0 + sqrt(3.0) / p.sphere_div) * (1.0 + 0x1p-16);
        p.cg_base = total;
        total += (int64_t)p.cg_n[0] * p.cg_n[1] * p.cg_n[2];
      }
      if (total <= cell_cap) {
        cg_out[0] = total;
        fit = true;
      } else {
        scale *= cbrt((double)total / (double)cell_cap) * 1.05;
      }
    }
    if (!fit) {
      for (int i = 0; i < ninst; i++) {
        P[i].cg_n[0] = P[i].cg_n[1] = P[i].cg_n[2] = 1;
        P[i].cg_base = i;
      }
      cg_out[0] = ninst;
    }
  }
  for (int i = 0; i < ninst; i++) {
    long long *a = acc + 8 * i;
    for (int k = 0; k < 3; k++) {
      a[k] = 0x7fffffffffffffffll;
      a[3 + k] = (long long)0x8000000000000000ull;
    }
    a[6] = (long long)0x8000000000000000ull;
    a[7] = 0x7fffffffffffffffll;
  }
}

// ---------------------------------------------------------------------- reorder -----

// Key of each clump for the spatial reorder: the sphere cell of its centre.
// Storage order of the spatial reorder: the instance's sphere cells taken in tiles of
// TILE^3 cells (tiles in row-major order, cells row-major inside a tile, tiles at the upper
// edges truncated).  A bijection on the instance's cell range, so the counting sort over
// the binning's cell counts applies.  Consecutive clumps then fill compact blocks instead of
// rows of the box, so the wavefronts away from the walls skip the wall modules entirely.
template <int TILE>
__device__ __forceinline__ uint32_t tiled_cell_key(const InstParams &p, d3 q) {
  const int nx = p.sg_n[0], ny = p.sg_n[1], nz = p.sg_n[2];
  const int ix = grid_index(q.x, p.sg_lo[0], p.sg_inv_h, nx);
  const int iy = grid_index(q.y, p.sg_lo[1], p.sg_inv_h, ny);
  const int iz = grid_index(q.z, p.sg_lo[2], p.sg_inv_h, nz);
  const int tx = ix / TILE, ty = iy / TILE, tz = iz / TILE;
  const int wx = min(TILE, nx - tx * TILE), hy = min(TILE, ny - ty * TILE);
  const int hz = min(TILE, nz - tz * TILE);
  // cells of the full tile slabs below, the full tile rows before in this slab, the tiles
  // before in this row, then the cell inside its tile
  const int64_t base = (int64_t)tz * TILE * nx * ny + (int64_t)ty * TILE * nx * hz +
                       (int64_t)tx * TILE * hy * hz;
  return (uint32_t)(p.sg_base + base + (ix - tx * TILE) +
                    (int64_t)wx * ((iy - ty * TILE) + (int64_t)hy * (iz - tz * TILE)));
}
constexpr int kReorderTile = 16;

__global__ void k_reorder_key(const double *__restrict__ st, const int32_t *__restrict__ inst_of,
                              const InstParams *__restrict__ P, uint32_t *__restrict__ counts,
                              uint32_t *__restrict__ ckey, uint32_t *__restrict__ crank,
                              int64_t n, int64_t cap) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const InstParams &p = P[inst_of[c]];
  const uint32_t key = tiled_cell_key<kReorderTile>(
      p, make_d3(st[S_RX * cap + c], st[S_RY * cap + c], st[S_RZ * cap + c]));
  ckey[c] = key;
  crank[c] = atomicAdd(&counts[key], 1u);
}

// new index = starts[key] + rank; gathers every per-clump array into the new order.
__global__ void k_reorder_move(const uint32_t *__restrict__ ckey, const uint32_t *__restrict__ crank,
                               const uint32_t *__restrict__ starts, const double *__restrict__ st,
                               double *__restrict__ st2, const double *__restrict__ sta,
                               double *__restrict__ sta2, const int32_t *__restrict__ orig,
                               int32_t *__restrict__ orig2, const int32_t *__restrict__ inst_of,
                               int32_t *__restrict__ inst2, const InstParams *__restrict__ P,
                               int64_t *__restrict__ inv, int64_t n, int64_t cap) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int64_t d = (int64_t)starts[ckey[c]] + crank[c];
  for (int f = 0; f < S_NFIELDS; f++) st2[f * cap + d] = st[f * cap + c];
  for (int f = 0; f < C_NFIELDS; f++) sta2[f * cap + d] = sta[f * cap + c];
  orig2[d] = orig[c];
  inst2[d] = inst_of[c];
  inv[P[inst_of[c]].begin + orig[c]] = d;
}

// ----------------------------------------------------------- export / debug kernels ---

// Reference AoS records (cabecera.hu:77-98) from the device state, in reference order.
// Sphere centres use the pre-freeze radius (stale centres of a frozen clump); rrot_* are
// the rotate_vectors.cu:28-46 images of the centres.
__global__ void k_export_aos(const double *__restrict__ st, const double *__restrict__ sta,
                             const int32_t *__restrict__ orig, const int32_t *__restrict__ inst_of,
                             const InstParams *__restrict__ P, int32_t inst, double *__restrict__ out,
                             double *__restrict__ aa, double *__restrict__ fo,
                             const double *__restrict__ fin, int64_t n, int64_t cap) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n || inst_of[c] != inst) return;
  const InstParams &p = P[inst];
  const int64_t o = orig[c];
  double *r = out + o * 51;
  const d3 rr = make_d3(st[S_RX * cap + c], st[S_RY * cap + c], st[S_RZ * cap + c]);
  double R[9];
  for (int k = 0; k < 9; k++) R[k] = st[(S_R00 + k) * cap + c];
  d3 sph[4];
  sphere_centres(rr, R, sta[C_RAD0 * cap + c] / p.sphere_div, sph);
  r[0] = sta[C_RAD * cap + c];
  r[1] = sta[C_MASS * cap + c];
  r[2] = sta[C_RHO * cap + c];
  r[3] = sta[C_IX * cap + c];
  r[4] = sta[C_IY * cap + c];
  r[5] = sta[C_IZ * cap + c];
  for (int s = 0; s < 4; s++) {
    r[6 + 3 * s] = sph[s].x;
    r[7 + 3 * s] = sph[s].y;
    r[8 + 3 * s] = sph[s].z;
    const d3 q = to_slope(sph[s], p.cost, p.sint);
    r[18 + 3 * s] = q.x;
    r[19 + 3 * s] = q.y;
    r[20 + 3 * s] = q.z;
  }
  r[30] = rr.x;
  r[31] = rr.y;
  r[32] = rr.z;
  const d3 q = to_slope(rr, p.cost, p.sint);
  r[33] = q.x;
  r[34] = q.y;
  r[35] = q.z;
  for (int k = 0; k < 3; k++) r[36 + k] = st[(S_VX + k) * cap + c];
  for (int k = 0; k < 3; k++) r[39 + k] = st[(S_WX + k) * cap + c];
  for (int k = 0; k < 9; k++) r[42 + k] = R[k];
  if (aa)
    for (int k = 0; k < 6; k++) aa[o * 6 + k] = st[(S_AX + k) * cap + c];
  if (fo)
    for (int k = 0; k < 6; k++) fo[o * 6 + k] = fin[k * cap + c];
}

// Inverse of k_export_aos for uploads (sphere centres and rrot are derived, not stored).
// extra holds 8 doubles per clump: a_i[3], alpha_i[3], pre-freeze radius, unused.  Clears
// P->mass_closed (set beforehand to 1 + clump_mode, or 0) when a clump's mass or inertia
// differs from the closed forms of its pre-freeze radius and density.
__global__ void k_import_aos(double *__restrict__ st, double *__restrict__ sta,
                             const int64_t *__restrict__ inv, int64_t begin,
                             const double *__restrict__ in, const double *__restrict__ extra,
                             InstParams *__restrict__ P, int64_t count, int64_t cap) {
  int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= count) return;
  const int64_t c = inv[begin + o];
  const double *r = in + o * 51;
  const double *x = extra + o * 8;
  if (P->mass_closed) {
    double m, ic;
    clump_mass_inertia(P->mass_closed - 1, x[6], r[2], P->sphere_div, m, ic);
    if (!(r[1] == m && r[3] == ic && r[4] == ic && r[5] == ic)) atomicAnd(&P->mass_closed, 0);
  }
  sta[C_RAD * cap + c] = r[0];
  sta[C_MASS * cap + c] = r[1];
  sta[C_RHO * cap + c] = r[2];
  sta[C_IX * cap + c] = r[3];
  sta[C_IY * cap + c] = r[4];
  sta[C_IZ * cap + c] = r[5];
  sta[C_RAD0 * cap + c] = x[6];
  for (int k = 0; k < 3; k++) st[(S_RX + k) * cap + c] = r[30 + k];
  for (int k = 0; k < 3; k++) st[(S_VX + k) * cap + c] = r[36 + k];
  for (int k = 0; k < 3; k++) st[(S_WX + k) * cap + c] = r[39 + k];
  for (int k = 0; k < 9; k++) st[(S_R00 + k) * cap + c] = r[42 + k];
  for (int k = 0; k < 6; k++) st[(S_AX + k) * cap + c] = x[k];
}

// Reference clump-cell index per clump (cells.cu:69-91), -1 for clumps cell_locate skips.
__global__ void k_export_cells(const double *__restrict__ st, const int32_t *__restrict__ orig,
                               const int32_t *__restrict__ inst_of, const InstParams *__restrict__ P,
                               int32_t inst, int64_t *__restrict__ out, int64_t n, int64_t cap) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n || inst_of[c] != inst) return;
  const InstParams &p = P[inst];
  const d3 r = make_d3(st[S_RX * cap + c], st[S_RY * cap + c], st[S_RZ * cap + c]);
  int64_t v = -1;
  if (!frozen(p.time, r.x, p.side_x)) {
    int64_t ii = (int64_t)(p.cs_inv[0] * (r.x - p.rr_min[0]));
    if (ii == p.ncell[0]) ii--;
    int64_t jj = (int64_t)(p.cs_inv[1] * (r.y - p.rr_min[1]));
    if (jj == p.ncell[1]) jj--;
    int64_t kk = (int64_t)(p.cs_inv[2] * (r.z - p.rr_min[2]));
    if (kk == p.ncell[2]) kk--;
    v = (ii < 0 || jj < 0 || kk < 0) ? -2 : ii + p.ncell[0] * (jj + kk * p.ncell[1]);
  }
  out[orig[c]] = v;
}

// Directed contacts of the current state through the sphere grid (debug export).
__global__ void k_export_contacts(const double *__restrict__ st, const double *__restrict__ sta,
                                  const double *__restrict__ stage, const int32_t *__restrict__ orig,
                                  const int32_t *__restrict__ inst_of, const InstParams *__restrict__ P,
                                  int32_t inst, const uint32_t *__restrict__ starts,
                                  const double4 *__restrict__ spos, const uint32_t *__restrict__ sid,
                                  int64_t *__restrict__ out, unsigned long long *__restrict__ count,
                                  int64_t capn, int64_t n, int64_t cap) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n || inst_of[c] != inst) return;
  const InstParams &p = P[inst];
  const double *gl = stage + c * kStageDoubles;
  const d3 rr = make_d3(gl[G_RR], gl[G_RR + 1], gl[G_RR + 2]);
  if (frozen(p.time, rr.x, p.side_x)) return;
  const double rad = gl[G_RAD], rho = sta[C_RHO * cap + c];
  const d3 vv = make_d3(gl[G_VV], gl[G_VV + 1], gl[G_VV + 2]);
  const d3 ww = make_d3(gl[G_WW], gl[G_WW + 1], gl[G_WW + 2]);
  const double massl = kSphFact * rho * rad * rad * rad;
  const int nx = p.sg_n[0], ny = p.sg_n[1], nz = p.sg_n[2];
  for (int s = 0; s < 4; s++) {
    const d3 q = rec_sphere(gl, rad / p.sphere_div, s);
    const int ix = grid_index(q.x, p.sg_lo[0], p.sg_inv_h, nx);
    const int iy = grid_index(q.y, p.sg_lo[1], p.sg_inv_h, ny);
    const int iz = grid_index(q.z, p.sg_lo[2], p.sg_inv_h, nz);
    const int x0 = ix > 0 ? ix - 1 : 0, x1 = ix + 1 < nx ? ix + 1 : nx - 1;
    for (int z = iz - 1; z <= iz + 1; z++) {
      if (z < 0 || z >= nz) continue;
      for (int y = iy - 1; y <= iy + 1; y++) {
        if (y < 0 || y >= ny) continue;
        const int64_t row = p.sg_base + (int64_t)nx * (y + (int64_t)ny * z);
        for (uint32_t j = starts[row + x0]; j < starts[row + x1 + 1]; j++) {
          const uint32_t id = sid[j];
          const int64_t cm = (int64_t)(id >> 2);
          if (cm == c) continue;
          const double4 o = spos[j];
          const double sep = rad + o.w, sep2 = sep * sep;
          const double dx = o.x - q.x, dy = o.y - q.y, dz = o.z - q.z;
          const double d2 = dx * dx + dy * dy + dz * dz;
          if (!(d2 < sep2)) continue;
          const double *g = stage + cm * kStageDoubles;
          const double rad_2 = g[G_RAD], rhom = sta[C_RHO * cap + cm];
          const double massm = kSphFact * rhom * rad_2 * rad_2 * rad_2;
          const double meff = 0.5 * (massm + massl);
          const double km = 0.5 * p.kappa_pp * meff, gm = 0.5 * p.gamma_pp * meff;
          const double dist = sqrt(d2);
          const d3 nvec = make_d3(dx / dist, dy / dist, dz / dist);
          const d3 rrc = make_d3(q.x + rad * nvec.x, q.y + rad * nvec.y, q.z + rad * nvec.z);
          const d3 rrm = make_d3(g[G_RR], g[G_RR + 1], g[G_RR + 2]);
          const d3 vl = point_vel(vv, ww, rrc, rr);
          const d3 vm = point_vel(make_d3(g[G_VV], g[G_VV + 1], g[G_VV + 2]),
                                  make_d3(g[G_WW], g[G_WW + 1], g[G_WW + 2]), rrc, rrm);
          Contact cl;
          const bool rep =
              contact_normal(cl, make_d3(vm.x - vl.x, vm.y - vl.y, vm.z - vl.z), nvec,
                             sep - dist, km, gm) > 0.0;
          const unsigned long long k = atomicAdd(count, 1ull);
          if ((int64_t)k < capn) {
            out[4 * k + 0] = orig[c];
            out[4 * k + 1] = orig[cm];
            out[4 * k + 2] = ((int64_t)s) | ((int64_t)(id & 3) << 8);
            out[4 * k + 3] = rep ? 1 : 0;
          }
        }
      }
    }
  }
}

// ------------------------------------------------------------------- observables ----

constexpr int kObsBlock = 256;
constexpr int kObsScalars = 9;  // |v|, |w|, m x, m z, m x', max x', m_live, live, m
constexpr int kMaxBins = 1024;

__device__ __forceinline__ double tree_sum(double *red, double v) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kObsBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double tree_max(double *red, double v) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kObsBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double gauss(double u, double sigma) {
  const double k = 0.3989422804014327;  // 1/sqrt(2 pi)
  return k / sigma * exp(-0.5 * (u / sigma) * (u / sigma));
}

// Per-block partial sums over reference indices [b*256, (b+1)*256) of one instance in a
// fixed tree order, so the result does not depend on the storage order.  partial holds
// kObsScalars + nbins doubles per block.
__global__ __launch_bounds__(kObsBlock) void k_observe(
    const double *__restrict__ st, const double *__restrict__ sta, const int64_t *__restrict__ inv,
    const InstParams *__restrict__ P, int32_t inst, int64_t count, int64_t cap, int nbins,
    double x0, double bw, double sigma, double width, double *__restrict__ partial) {
  __shared__ double red[kObsBlock];
  const InstParams &p = P[inst];
  const int64_t o = (int64_t)blockIdx.x * kObsBlock + threadIdx.x;
  double sv = 0.0, sw = 0.0, mx = 0.0, mz = 0.0, mxs = 0.0, xmax = -1.0e300, ml = 0.0, nl = 0.0;
  double mall = 0.0;
  double xs = 0.0, vol = 0.0;
  bool live = false;
  if (o < count) {
    const int64_t c = inv[p.begin + o];
    const double x = st[S_RX * cap + c], z = st[S_RZ * cap + c];
    const double vx = st[S_VX * cap + c], vy = st[S_VY * cap + c], vz = st[S_VZ * cap + c];
    const double wx = st[S_WX * cap + c], wy = st[S_WY * cap + c], wz = st[S_WZ * cap + c];
    const double m = sta[C_MASS * cap + c];
    double vm2 = vx * vx;  // avalanche_hst.cu:551-558
    vm2 += vy * vy;
    vm2 += vz * vz;
    sv = sqrt(vm2);
    double wm2 = wx * wx;
    wm2 += wy * wy;
    wm2 += wz * wz;
    sw = sqrt(wm2);
    mx = m * x;  // avalanche_hst.cu:822-823
    mz = m * z;
    mall = m;
    live = !frozen(p.time, x, p.side_x);
    if (live) {
      xs = x * p.cost - z * p.sint;
      mxs = m * xs;
      xmax = xs;
      ml = m;
      nl = 1.0;
      vol = m / sta[C_RHO * cap + c];
    }
  }
  double *out = partial + (size_t)blockIdx.x * (kObsScalars + nbins);
  const double r0 = tree_sum(red, sv), r1 = tree_sum(red, sw), r2 = tree_sum(red, mx);
  const double r3 = tree_sum(red, mz), r4 = tree_sum(red, mxs), r5 = tree_max(red, xmax);
  const double r6 = tree_sum(red, ml), r7 = tree_sum(red, nl), r8 = tree_sum(red, mall);
  if (threadIdx.x == 0) {
    out[0] = r0; out[1] = r1; out[2] = r2; out[3] = r3;
    out[4] = r4; out[5] = r5; out[6] = r6; out[7] = r7;
    out[8] = r8;
  }
  for (int b = 0; b < nbins; b++) {
    const double cb = x0 + (b + 0.5) * bw;
    const double u = xs - cb;
    const double contrib = (live && fabs(u) < 8.0 * sigma) ? vol * gauss(u, sigma) / width : 0.0;
    const double rb = tree_sum(red, contrib);
    if (threadIdx.x == 0) out[kObsScalars + b] = rb;
  }
}

// Sums the block partials in block order (single thread).
__global__ void k_observe_final(const double *__restrict__ partial, int64_t nblocks, int nbins,
                                double *__restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int w = kObsScalars + nbins;
  for (int k = 0; k < w; k++) {
    double acc = (k == 5) ? -1.0e300 : 0.0;
    for (int64_t b = 0; b < nblocks; b++) {
      const double v = partial[b * w + k];
      acc = (k == 5) ? fmax(acc, v) : acc + v;
    }
    out[k] = acc;
  }
}

}  // namespace avf
