// dem_kernels.h -- device data layout and kernel interface of the gfx950 DEM step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dem_math.h"

namespace avf {

// Per-instance parameters, resident in device memory and read by every kernel.  time and
// cc advance on the device, so a captured graph replays correctly.
struct InstParams {
  // system_par fields the ops read (cabecera.hu:34-66)
  double dt, hdt, time, side_x, side_y;
  double sint, cost;
  double kappa_pp, gamma_pp, mu_pp, fr_pp;
  double kappa_pw, gamma_pw, mu_pw, fr_pw;
  double dy2, sinfi, cosfi, tanfi;
  // geometry (cabecera.hu:14-19 as runtime values) and the deposit plane
  double hop_height, hop_width, yfunnel, lfunnel, yhop_l, yhop_r;
  double grav, drag, sphere_div;
  double dep_sin, dep_cos, dep_x0, dep_x1, dep_z0;
  // clump-cell regrid (cell_dim.cu / avalanche_hst.cu:311-333)
  double rmax, gyr_max, cell_dist, regrid_ddt_steps;
  double rr_min[3], cell_side[3], cs_inv[3], rr_max[3], length[3];
  int64_t ncell[3];
  // sphere-cell grid of this instance (counting-sort binning)
  double sg_lo[3], sg_inv_h, sg_h, sg_reach, sg_rsph;
  int32_t sg_n[3], pad0;
  int64_t sg_base;
  int64_t sg_bbase;  // first brick of this instance in the contact sweep's brick range
  // clump-cell grid of this instance (per-step binning by clump: one key per clump centre,
  // cell side >= the largest clump's bounding diameter); cg_mh: the float test's margin
  // (length), cg_bfac: bounding radius / sphere radius, rounded up
  double cg_lo[3], cg_inv_h, cg_h, cg_mh, cg_bfac;
  int32_t cg_n[3], pad1;
  int64_t cg_base;
  // storage range and bookkeeping
  int64_t begin, count;
  int64_t cc;
  int32_t regrid_interval, inter_flag, phase, modules_fill;
  // 1 + clump_mode when every clump's stored mass and inertia equal clump_mass_inertia() of
  // its radius and density bit for bit (checked on upload) and the context asked for it:
  // the step then recomputes them instead of loading them; 0 otherwise
  int32_t modules_aval, mass_closed;
};

// Clump state fields (SoA, field-major: st[f * cap + c]).
enum StateField {
  S_RX, S_RY, S_RZ, S_VX, S_VY, S_VZ, S_WX, S_WY, S_WZ,
  S_R00, S_R01, S_R02, S_R10, S_R11, S_R12, S_R20, S_R21, S_R22,
  S_AX, S_AY, S_AZ, S_GX, S_GY, S_GZ, S_NFIELDS
};
// Static per-clump fields (radius changes only when a clump freezes).
enum StaticField { C_RAD, C_RAD0, C_RHO, C_MASS, C_IX, C_IY, C_IZ, C_NFIELDS };
// Force / torque of the last force stage.
enum ForceField { F_FX, F_FY, F_FZ, F_TX, F_TY, F_TZ, F_NFIELDS };

// 176-B record per clump written by the predict pass and gathered by neighbours:
// rr[3], rad, vv[3], ww[3], R rows[9], a[3] (the acceleration the step's finish uses).
// With R and a it holds the whole carried state except the angular acceleration, so inside
// a chunk of list steps a clump can live in its record alone.  A clump's density and
// reference index, which no step changes, are read from sta and orig.
// Sphere centres are not stored: whoever needs one recomputes rr + pos R (+-1, +-1, +-1)
// with sphere_of, bit-identical to the predict's, pos = (radius before freezing) / div.
constexpr int kStageDoubles = 22;
enum StageSlot { G_RR = 0, G_RAD = 3, G_VV = 4, G_WW = 7, G_R = 10, G_AA = 19 };

constexpr uint32_t kInvalidKey = 0xffffffffu;
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

// Device-side diagnostics (avf_diag).
struct DevDiag {
  unsigned long long cell_below, cell_beyond, neigh_overflow, max_neighbours, max_cell;
  unsigned long long sweep_global, sweep_fp64;  // bricks swept from global memory, FP64 pairs
  unsigned long long bin_broken;  // scattered entries beyond the binned count (never expected)
};

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on XCD b % 8),
// each with its own L2.  This maps the hardware workgroup index to a logical one so that
// every XCD processes one contiguous range of logical workgroups: spatially close clumps
// (the storage order is spatial) then share one L2 instead of being fetched by all eight.
// A bijection on [0, nb) for any nb.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
  const unsigned q = nb / kXcds, r = nb % kXcds, x = b % kXcds, k = b / kXcds;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

// The same idea at a finer grain for kernels whose work per workgroup varies with the local
// density (k_force): XCD x processes runs of L consecutive logical workgroups, runs x, x + 8,
// x + 16, ... in turn, so dense regions are shared by all XCDs while neighbouring workgroups
// still share one L2.  A bijection on [0, nb) (the last nb mod 8L workgroups map to
// themselves).
__device__ __forceinline__ unsigned xcd_block_runs(unsigned b, unsigned nb, unsigned L) {
  const unsigned span = kXcds * L, full = nb / span * span;
  if (b >= full) return b;
  const unsigned x = b % kXcds, k = b / kXcds;
  return ((k / L) * kXcds + x) * L + (k % L);
}

// Running maximum of a diagnostic counter without contending on one address: a lane only
// issues the atomic when its value exceeds what the counter already holds (after the first
// steps almost no lane does).
__device__ __forceinline__ void diag_max(unsigned long long *addr, unsigned long long v) {
  if (v > __atomic_load_n(addr, __ATOMIC_RELAXED)) atomicMax(addr, v);
}

// Order-preserving int64 image of a double (for integer atomic min / max).
__device__ __forceinline__ long long order_key(double v) {
  long long b = __double_as_longlong(v);
  return b >= 0 ? b : (b ^ 0x7fffffffffffffffll);
}
__device__ __forceinline__ double order_val(long long k) {
  return __longlong_as_double(k >= 0 ? k : (k ^ 0x7fffffffffffffffll));
}

}  // namespace avf
