// dem_math.h -- per-contact arithmetic of the DEM step, shared by the gfx950 kernels.
//
// Every expression keeps the operation order of the reference's host ops (the kernels are
// compiled with -ffp-contract=off and IEEE division / square root), so forces, torques and
// trajectories are bit-identical to the host path on the same inputs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace avf {

// Literals of cabecera.hu:7-12, as written (SPH_FACT is 1 ulp above 4*pi/3).
constexpr double kSphFact = 4.188790204786391;

struct d3 {
  double x, y, z;
};

__device__ __forceinline__ d3 make_d3(double x, double y, double z) {
  d3 r;
  r.x = x;
  r.y = y;
  r.z = z;
  return r;
}

__device__ __forceinline__ bool frozen(double time, double x, double side_x) {
  // "time > 0.0 && rr.x > side.x" (init_Verlet.cu:29, cells.cu:65, force_part.cu:47, ...)
  return time > 0.0 && x > side_x;
}

// Contact-point velocity v + w x (rrc - rr), force_part.cu:175-181.
__device__ __forceinline__ d3 point_vel(d3 vv, d3 ww, d3 rrc, d3 rr) {
  d3 v;
  v.x = vv.x + ww.y * (rrc.z - rr.z) - ww.z * (rrc.y - rr.y);
  v.y = vv.y + ww.z * (rrc.x - rr.x) - ww.x * (rrc.z - rr.z);
  v.z = vv.z + ww.x * (rrc.y - rr.y) - ww.y * (rrc.x - rr.x);
  return v;
}

// tq += (rrc - rr) x F, ff += F (force_part.cu:245-251).
__device__ __forceinline__ void accumulate(d3 &ff, d3 &tq, d3 rrc, d3 rr, d3 f) {
  tq.x += (rrc.y - rr.y) * f.z - (rrc.z - rr.z) * f.y;
  tq.y += (rrc.z - rr.z) * f.x - (rrc.x - rr.x) * f.z;
  tq.z += (rrc.x - rr.x) * f.y - (rrc.y - rr.y) * f.x;
  ff.x += f.x;
  ff.y += f.y;
  ff.z += f.z;
}

// Relative-velocity decomposition and normal spring-dashpot force of one contact
// (force_part.cu:188-208; the wall modules repeat it verbatim).
struct Contact {
  d3 nvec, dvvct;
  double dvvct_mag, dvvct_sig, normal;
};

__device__ __forceinline__ double contact_normal(Contact &c, d3 dvvc, d3 nvec, double comp,
                                                 double km, double gm) {
  double dvvcn_mag, comp_dot;
  d3 dvvcn, tvec;
  c.nvec = nvec;
  dvvcn_mag = dvvc.x * nvec.x + dvvc.y * nvec.y + dvvc.z * nvec.z;
  dvvcn.x = dvvcn_mag * nvec.x;
  dvvcn.y = dvvcn_mag * nvec.y;
  dvvcn.z = dvvcn_mag * nvec.z;
  c.dvvct.x = dvvc.x - dvvcn.x;
  c.dvvct.y = dvvc.y - dvvcn.y;
  c.dvvct.z = dvvc.z - dvvcn.z;
  c.dvvct_mag = sqrt(c.dvvct.x * c.dvvct.x + c.dvvct.y * c.dvvct.y + c.dvvct.z * c.dvvct.z);
  // 0/0 = NaN for a vanishing tangential velocity; the dvvct_mag > 0 guard of the
  // friction law keeps it out of the force (force_part.cu:199-222).
  tvec.x = c.dvvct.x / c.dvvct_mag;
  tvec.y = c.dvvct.y / c.dvvct_mag;
  tvec.z = c.dvvct.z / c.dvvct_mag;
  c.dvvct_sig = dvvc.x * tvec.x + dvvc.y * tvec.y + dvvc.z * tvec.z;
  comp_dot = -dvvcn_mag;
  c.normal = km * comp + gm * comp_dot;
  return c.normal;
}

// Sphere force -N n plus the one-step tangential spring-dashpot capped by a regularised
// Coulomb law.  pp_rule: slide iff |F_sdt| > mu N (host pp rule, force_part.cu:224);
// otherwise the wall rule, slide iff |F_sdt| > mu N or xi == 0 (force_hop.cu:207).
__device__ __forceinline__ d3 contact_friction(const Contact &c, double mu, double cns_1,
                                               double dt, double mass_f, bool pp_rule) {
  double normal = c.normal, xi, xi_dot, ff_sdt, aux_fric, t0;
  d3 f;
  f.x = -normal * c.nvec.x;
  f.y = -normal * c.nvec.y;
  f.z = -normal * c.nvec.z;
  xi_dot = c.dvvct_sig;
  xi = dt * xi_dot;
  if (cns_1 * normal < fabs(xi)) xi = 0.0;
  xi_dot = xi / dt;
  ff_sdt = (1.0e12 * mass_f * xi + 2.0e6 * mass_f * xi_dot);
  if (c.dvvct_mag > 0.0) {
    bool slide = pp_rule ? (fabs(ff_sdt) > mu * normal)
                         : (fabs(ff_sdt) > mu * normal || xi == 0.0);
    if (slide) {
      aux_fric = 0.8 * mu * normal / c.dvvct_mag;
      t0 = mass_f / aux_fric;
      aux_fric /= (1.0 + (dt / t0) * (dt / t0));
      f.x += aux_fric * c.dvvct.x;
      f.y += aux_fric * c.dvvct.y;
      f.z += aux_fric * c.dvvct.z;
    } else {
      if (c.dvvct.x != 0.0) f.x += ff_sdt * c.dvvct.x / c.dvvct_mag;
      if (c.dvvct.y != 0.0) f.y += ff_sdt * c.dvvct.y / c.dvvct_mag;
      if (c.dvvct.z != 0.0) f.z += ff_sdt * c.dvvct.z / c.dvvct_mag;
    }
  }
  return f;
}

// Wall contact: relative velocity -v_c.  nan_applies mirrors force_funnel.cu:354
// (corners skip only on `normal <= 0.0`); the other wall modules apply on `normal > 0.0`.
__device__ __forceinline__ void wall_contact(d3 rr, d3 vv, d3 ww, d3 rrc, d3 nvec, double comp,
                                             double km, double gm, double mu, double cns_1,
                                             double dt, double mass, bool nan_applies, d3 &ff,
                                             d3 &tq) {
  d3 vvc = point_vel(vv, ww, rrc, rr), dvvc;
  dvvc.x = -vvc.x;
  dvvc.y = -vvc.y;
  dvvc.z = -vvc.z;
  Contact c;
  double normal = contact_normal(c, dvvc, nvec, comp, km, gm);
  if (nan_applies ? (normal <= 0.0) : !(normal > 0.0)) return;
  d3 f = contact_friction(c, mu, cns_1, dt, mass, false);
  accumulate(ff, tq, rrc, rr, f);
}

// Slope frame -> lab frame (force_funnel.cu:88-93).
__device__ __forceinline__ d3 to_lab(d3 v, double cost, double sint) {
  d3 r;
  r.x = v.x * cost + v.z * sint;
  r.y = v.y;
  r.z = v.z * cost - v.x * sint;
  return r;
}

// Lab frame -> slope frame (rotate_vectors.cu:28-30).
__device__ __forceinline__ d3 to_slope(d3 v, double cost, double sint) {
  d3 r;
  r.x = v.x * cost - v.z * sint;
  r.y = v.y;
  r.z = v.x * sint + v.z * cost;
  return r;
}

}  // namespace avf
