/*
 * dem_oracle.c -- CPU oracle for the DEM step (TEST INFRASTRUCTURE ONLY; see dem_oracle.h).
 *
 * Restates the reference's host ops with the same operation order and literals.  Build:
 * gcc -O2, no -march, no fast-math, -ffp-contract=off (oracle/Makefile), mirroring the
 * reference's `nvcc -O2` host compile (MakeAvalancheHst.mk:45-103).
 */
#include "dem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline int orc_frozen(double time, double x, double side_x) {
  /* "time > 0.0 && rr.x > side.x": init_Verlet.cu:29, cells.cu:65, force_part.cu:47 ... */
  return time > 0.0 && x > side_x;
}

static inline avf_double3 orc_sphere(const avf_particle *p, int s) {
  switch (s) {
    case 0: return p->rr_a;
    case 1: return p->rr_b;
    case 2: return p->rr_c;
    default: return p->rr_d;
  }
}

static inline avf_double3 orc_sphere_rot(const avf_particle *p, int s) {
  switch (s) {
    case 0: return p->rrot_a;
    case 1: return p->rrot_b;
    case 2: return p->rrot_c;
    default: return p->rrot_d;
  }
}

/* One contact's relative-velocity decomposition and normal force, shared by every module
 * (force_part.cu:188-208, force_hop.cu:170-189, force_funnel.cu:137-156, force_funnel.cu:
 * 332-351, force_walls.cu:123-142, force_flume.cu:106-125, force_dep.cu). */
typedef struct orc_cl {
  avf_double3 nvec, dvvct;
  double dvvct_mag, dvvct_sig, normal;
} orc_cl;

static double orc_normal(orc_cl *c, avf_double3 dvvc, avf_double3 nvec, double comp, double km,
                         double gm) {
  double dvvcn_mag, comp_dot;
  avf_double3 dvvcn, tvec;
  c->nvec = nvec;
  dvvcn_mag = dvvc.x * nvec.x + dvvc.y * nvec.y + dvvc.z * nvec.z;
  dvvcn.x = dvvcn_mag * nvec.x;
  dvvcn.y = dvvcn_mag * nvec.y;
  dvvcn.z = dvvcn_mag * nvec.z;
  c->dvvct.x = dvvc.x - dvvcn.x;
  c->dvvct.y = dvvc.y - dvvcn.y;
  c->dvvct.z = dvvc.z - dvvcn.z;
  c->dvvct_mag = sqrt(c->dvvct.x * c->dvvct.x + c->dvvct.y * c->dvvct.y + c->dvvct.z * c->dvvct.z);
  /* 0/0 = NaN when the tangential velocity vanishes; the dvvct_mag > 0 guard of the
   * friction law keeps it out of the force (force_part.cu:199-222). */
  tvec.x = c->dvvct.x / c->dvvct_mag;
  tvec.y = c->dvvct.y / c->dvvct_mag;
  tvec.z = c->dvvct.z / c->dvvct_mag;
  c->dvvct_sig = dvvc.x * tvec.x + dvvc.y * tvec.y + dvvc.z * tvec.z;
  comp_dot = -dvvcn_mag;
  c->normal = km * comp + gm * comp_dot;
  return c->normal;
}

/* Sphere force -N n plus the one-step tangential spring-dashpot with the regularised
 * Coulomb cap: force_part.cu:211-242 (pp_rule = 1: slide iff |F_sdt| > mu N,
 * force_part.cu:224) or the wall rule (pp_rule = 0: slide iff |F_sdt| > mu N or xi == 0,
 * force_hop.cu:207).  mass_f is the mass of the friction law. */
static avf_double3 orc_friction(const orc_cl *c, double mu, double cns_1, double dt,
                                double mass_f, int pp_rule) {
  double normal = c->normal, xi, xi_dot, ff_sdt, aux_fric, t0;
  avf_double3 f;
  f.x = -normal * c->nvec.x;
  f.y = -normal * c->nvec.y;
  f.z = -normal * c->nvec.z;
  xi_dot = c->dvvct_sig;
  xi = dt * xi_dot;
  if (cns_1 * normal < fabs(xi)) xi = 0.0;
  xi_dot = xi / dt;
  ff_sdt = (1.0e12 * mass_f * xi + 2.0e6 * mass_f * xi_dot);
  if (c->dvvct_mag > 0.0) {
    int slide = pp_rule ? (fabs(ff_sdt) > mu * normal) : (fabs(ff_sdt) > mu * normal || xi == 0.0);
    if (slide) {
      aux_fric = 0.8 * mu * normal / c->dvvct_mag;
      t0 = mass_f / aux_fric;
      aux_fric /= (1.0 + (dt / t0) * (dt / t0));
      f.x += aux_fric * c->dvvct.x;
      f.y += aux_fric * c->dvvct.y;
      f.z += aux_fric * c->dvvct.z;
    } else {
      if (c->dvvct.x != 0.0) f.x += ff_sdt * c->dvvct.x / c->dvvct_mag;
      if (c->dvvct.y != 0.0) f.y += ff_sdt * c->dvvct.y / c->dvvct_mag;
      if (c->dvvct.z != 0.0) f.z += ff_sdt * c->dvvct.z / c->dvvct_mag;
    }
  }
  return f;
}

/* Contact-point velocity v + w x (rrc - rr) (force_hop.cu:160-162). */
static inline avf_double3 orc_point_vel(avf_double3 vv, avf_double3 ww, avf_double3 rrc,
                                        avf_double3 rr) {
  avf_double3 v;
  v.x = vv.x + ww.y * (rrc.z - rr.z) - ww.z * (rrc.y - rr.y);
  v.y = vv.y + ww.z * (rrc.x - rr.x) - ww.x * (rrc.z - rr.z);
  v.z = vv.z + ww.x * (rrc.y - rr.y) - ww.y * (rrc.x - rr.x);
  return v;
}

/* tq += (rrc - rr) x F; ff += F (force_hop.cu:228-234). */
static inline void orc_accumulate(avf_double3 *ff, avf_double3 *tq, avf_double3 rrc,
                                  avf_double3 rr, avf_double3 f) {
  tq->x += (rrc.y - rr.y) * f.z - (rrc.z - rr.z) * f.y;
  tq->y += (rrc.z - rr.z) * f.x - (rrc.x - rr.x) * f.z;
  tq->z += (rrc.x - rr.x) * f.y - (rrc.y - rr.y) * f.x;
  ff->x += f.x;
  ff->y += f.y;
  ff->z += f.z;
}

/* Wall contact: the relative velocity is minus the sphere's contact-point velocity.
 * nan_applies: force_fcorners skips only on `normal <= 0.0` (force_funnel.cu:354), so a NaN
 * normal still applies there; every other wall module applies on `normal > 0.0`. */
static inline void orc_wall_contact(const avf_particle *p, avf_double3 rrc, avf_double3 nvec,
                                    double comp, double km, double gm, double mu, double cns_1,
                                    double dt, double mass, int nan_applies, avf_double3 *ff,
                                    avf_double3 *tq) {
  avf_double3 vvc = orc_point_vel(p->vv, p->ww, rrc, p->rr), dvvc, f;
  orc_cl c;
  double normal;
  dvvc.x = -vvc.x;
  dvvc.y = -vvc.y;
  dvvc.z = -vvc.z;
  normal = orc_normal(&c, dvvc, nvec, comp, km, gm);
  if (nan_applies ? (normal <= 0.0) : !(normal > 0.0)) return;
  f = orc_friction(&c, mu, cns_1, dt, mass, 0);
  orc_accumulate(ff, tq, rrc, p->rr, f);
}

/* ------------------------------------------------------------------ clear_forces ----- */

/* clear_forces.cu:8-25 */
void orc_clear_forces(int64_t n, avf_double3 *force, avf_double3 *torque) {
  avf_double3 zero = {0.0, 0.0, 0.0};
  for (int64_t i = 0; i < n; i++) {
    force[i] = zero;
    torque[i] = zero;
  }
}

/* ------------------------------------------------------------------ cells ---------- */

/* cells.cu:11-22 */
void orc_clear_cell(const avf_cell *cell, int64_t *cell_vec) {
  for (int64_t i = 0; i < cell->ntot; i++) cell_vec[i] = -1;
}

/* cells.cu:26-37 */
void orc_clear_ocup(const avf_cell *cell, int32_t *nocup) {
  for (int64_t i = 0; i < cell->ncell3; i++) nocup[i] = 0;
}

/* cells.cu:40-111: clump centre -> cell (MULTIPLY by cs_inv), slot in ascending clump
 * order.  Index > ncell is not checked (it aliases into another cell, as in the reference);
 * the linear index is bounded here so the oracle itself stays in memory. */
int orc_cell_locate(const avf_cell *cell, const avf_system *sys, const avf_particle *pp,
                    int32_t *nocup, int64_t *cell_vec, int64_t *bad) {
  int64_t n = sys->npart;
  for (int64_t m = 0; m < n; m++) {
    avf_double3 rr = pp[m].rr;
    int64_t ii, jj, kk, idx;
    int32_t shift;
    if (orc_frozen(sys->time, rr.x, sys->side.x)) continue;
    ii = (int64_t)(cell->cs_inv.x * (rr.x - cell->rr_min.x));
    if (ii == cell->ncell.x) ii--;
    jj = (int64_t)(cell->cs_inv.y * (rr.y - cell->rr_min.y));
    if (jj == cell->ncell.y) jj--;
    kk = (int64_t)(cell->cs_inv.z * (rr.z - cell->rr_min.z));
    if (kk == cell->ncell.z) kk--;
    if (ii < 0 || jj < 0 || kk < 0) {
      if (bad) *bad = m;
      return AVF_ERR_CELL_NEG;
    }
    idx = ii + cell->ncell.x * (jj + kk * cell->ncell.y);
    if (idx >= cell->ncell3) { /* reference: out-of-bounds write; reported here */
      if (bad) *bad = m;
      return AVF_ERR_CELL_NEG;
    }
    shift = nocup[idx];
    if (shift == cell->ntags - 1) {
      if (bad) *bad = m;
      return AVF_ERR_CELL_FULL;
    }
    nocup[idx]++;
    cell_vec[idx * cell->ntags + (int64_t)shift] = m;
  }
  return 0;
}

/* cell_dim.cu:7-66 (ddt = ddt_steps*dt, 2000 in the reference). */
void orc_cell_dim(avf_double3 *rr_max, avf_double3 *rr_min, const avf_particle *pp,
                  const avf_system *sys, double ddt_steps) {
  int32_t nn = (int32_t)sys->npart; /* `int nn` in the reference */
  double rgyr = 0.5 * sys->gyr_max, dt = sys->dt, ddt = ddt_steps * dt, vxmax, vzmin;
  avf_double3 rmax, rmin;
  rmax.x = rmin.x = pp[0].rr.x;
  rmax.y = rmin.y = pp[0].rr.y;
  rmax.z = rmin.z = pp[0].rr.z;
  vxmax = pp[0].vv.x;
  vzmin = pp[0].vv.z;
  for (int32_t i = 1; i < nn; i++) {
    avf_double3 rr = pp[i].rr, vv = pp[i].vv;
    if (rr.x < rmin.x) rmin.x = rr.x;
    if (rr.x > rmax.x) rmax.x = rr.x;
    if (rr.y < rmin.y) rmin.y = rr.y;
    if (rr.y > rmax.y) rmax.y = rr.y;
    if (rr.z < rmin.z) rmin.z = rr.z;
    if (rr.z > rmax.z) rmax.z = rr.z;
    if (vv.x > vxmax) vxmax = vv.x;
    if (vv.z < vzmin) vzmin = vv.z;
  }
  rmax.x += 1.05 * rgyr + vxmax * ddt;
  rmin.x -= 1.05 * rgyr;
  rmax.y += 1.05 * rgyr;
  rmin.y -= 1.05 * rgyr;
  rmax.z += 1.05 * rgyr;
  rmin.z -= 1.05 * rgyr - vzmin * ddt;
  *rr_max = rmax;
  *rr_min = rmin;
}

/* avalanche_hst.cu:311-333 (same arithmetic at :480-499 and :723-742). */
void orc_cell_setup(avf_cell *cell, const avf_particle *pp, const avf_system *sys,
                    const avf_geometry *geo) {
  double dist = geo->cell_dist;
  orc_cell_dim(&cell->rr_max, &cell->rr_min, pp, sys, geo->regrid_ddt_steps);
  cell->length.x = cell->rr_max.x - cell->rr_min.x;
  cell->length.y = cell->rr_max.y - cell->rr_min.y;
  cell->length.z = cell->rr_max.z - cell->rr_min.z;
  cell->ncell.x = (int64_t)(cell->length.x / (dist * sys->rmax));
  cell->ncell.y = (int64_t)(cell->length.y / (dist * sys->rmax));
  cell->ncell.z = (int64_t)(cell->length.z / (dist * sys->rmax));
  cell->cell_side.x = cell->length.x / ((double)cell->ncell.x);
  cell->cell_side.y = cell->length.y / ((double)cell->ncell.y);
  cell->cell_side.z = cell->length.z / ((double)cell->ncell.z);
  cell->cs_inv.x = 1.0 / cell->cell_side.x;
  cell->cs_inv.y = 1.0 / cell->cell_side.y;
  cell->cs_inv.z = 1.0 / cell->cell_side.z;
  cell->ncell3 = cell->ncell.x * cell->ncell.y * cell->ncell.z;
  cell->ntot = cell->ncell3 * cell->ntags;
}

/* ------------------------------------------------------------------ force_part ------ */

/* force_part.cu:11-266: pairing sph_id (1..16) of sphere smod of l against sphere sdiv of
 * every clump m != l in the 27 clump cells around l (found by DIVISION by cell_side). */
void orc_force_part(avf_double3 *force, avf_double3 *torque, const int64_t *cell_vec,
                    const int32_t *nocup_vec, const avf_particle *pp, const avf_system *sys,
                    const avf_cell *cell, int sph_id, orc_contacts *rec) {
  int64_t nn = sys->npart, ntags = cell->ntags;
  double mu = sys->mu_pp, kappa = sys->kappa_pp, gamma = sys->gamma_pp, dt = sys->dt;
  double cns_1 = sys->fr_fact_pp, time = sys->time;
  avf_long3 ncell = cell->ncell;
  avf_double3 cell_side = cell->cell_side, rr_min = cell->rr_min;
  int smod = (sph_id - 1) % 4, sdiv = (sph_id - 1) / 4;

  for (int64_t ll = 0; ll < nn; ll++) {
    avf_double3 rrl = pp[ll].rr, vvl, wwl, ff, tq, rr_1;
    double rhol, rad_1, massl;
    int64_t ii, jj, kk;
    if (orc_frozen(time, rrl.x, sys->side.x)) continue;
    vvl = pp[ll].vv;
    wwl = pp[ll].ww;
    rhol = pp[ll].density;
    rad_1 = pp[ll].radius;
    ff = force[ll];
    tq = torque[ll];
    massl = ORC_SPH_FACT * rhol * rad_1 * rad_1 * rad_1;
    rr_1 = orc_sphere(&pp[ll], smod);

    ii = (int64_t)((rrl.x - rr_min.x) / cell_side.x);
    if (ii == ncell.x) ii--;
    jj = (int64_t)((rrl.y - rr_min.y) / cell_side.y);
    if (jj == ncell.y) jj--;
    kk = (int64_t)((rrl.z - rr_min.z) / cell_side.z);
    if (kk == ncell.z) kk--;

    for (int idel = -1; idel <= 1; idel++)
      for (int jdel = -1; jdel <= 1; jdel++)
        for (int kdel = -1; kdel <= 1; kdel++) {
          int64_t in = ii + idel, jn = jj + jdel, kn = kk + kdel, cind, t0, t1;
          if (!(in >= 0 && jn >= 0 && kn >= 0 && in < ncell.x && jn < ncell.y && kn < ncell.z))
            continue;
          cind = in + ncell.x * (jn + ncell.y * kn);
          t0 = cind * ntags;
          t1 = t0 + nocup_vec[cind];
          for (int64_t tag = t0; tag < t1; tag++) {
            int64_t mm = cell_vec[tag];
            avf_double3 rrm, vvm, wwm, rr_2, drr, nvec, rrc, vvlc, vvmc, dvvc, f;
            double rad_2, rhom, massm, meff, km, gm, sep, sep2, dist2, dist, comp;
            int rep;
            if (mm == ll) continue;
            rrm = pp[mm].rr;
            rad_2 = pp[mm].radius;
            vvm = pp[mm].vv;
            wwm = pp[mm].ww;
            rhom = pp[mm].density;
            massm = ORC_SPH_FACT * rhom * rad_2 * rad_2 * rad_2;
            meff = 0.5 * (massm + massl);
            km = 0.5 * kappa * meff;
            gm = 0.5 * gamma * meff;
            rr_2 = orc_sphere(&pp[mm], sdiv);

            sep = rad_1 + rad_2;
            sep2 = sep * sep;
            drr.x = rr_2.x - rr_1.x;
            dist2 = drr.x * drr.x;
            if (!(dist2 < sep2)) continue;
            drr.y = rr_2.y - rr_1.y;
            dist2 += drr.y * drr.y;
            if (!(dist2 < sep2)) continue;
            drr.z = rr_2.z - rr_1.z;
            dist2 += drr.z * drr.z;
            if (!(dist2 < sep2)) continue;

            dist = sqrt(dist2);
            comp = sep - dist;
            nvec.x = drr.x / dist;
            nvec.y = drr.y / dist;
            nvec.z = drr.z / dist;
            rrc.x = rr_1.x + rad_1 * nvec.x;
            rrc.y = rr_1.y + rad_1 * nvec.y;
            rrc.z = rr_1.z + rad_1 * nvec.z;
            vvlc = orc_point_vel(vvl, wwl, rrc, rrl);
            vvmc = orc_point_vel(vvm, wwm, rrc, rrm);
            dvvc.x = vvmc.x - vvlc.x;
            dvvc.y = vvmc.y - vvlc.y;
            dvvc.z = vvmc.z - vvlc.z;
            {
              orc_cl c;
              rep = orc_normal(&c, dvvc, nvec, comp, km, gm) > 0.0;
              if (rep) f = orc_friction(&c, mu, cns_1, dt, massl, 1);
            }
            if (rec) {
              if (rec->count < rec->cap) {
                avf_contact *c = &rec->buf[rec->count];
                c->l = ll;
                c->m = mm;
                c->sl = smod;
                c->sm = sdiv;
                c->repulsive = rep;
                c->reserved = 0;
              }
              rec->count++;
            }
            if (rep) orc_accumulate(&ff, &tq, rrc, rrl, f);
          }
        }
    force[ll] = ff;
    torque[ll] = tq;
  }
}

/* ------------------------------------------------------------------ boundaries ----- */

/* force_hop.cu:9-244: hopper box walls (nw = 0: x walls, 1: y walls, 2: bottom gate that
 * acts only while time < 0).  No frozen-clump skip.  If both walls of a pair touch, the
 * second one wins (touch_flag). */
void orc_force_hop(avf_double3 *force, avf_double3 *torque, const avf_system *sys,
                   const avf_particle *pp, const avf_geometry *geo, int nw) {
  int64_t nn = sys->npart;
  double mu = sys->mu_pw, kappa = sys->kappa_pw, gamma = sys->gamma_pw, dt = sys->dt;
  double cns_1 = sys->fr_fact_pw, time = sys->time;
  double xfront = geo->hop_width, xrear = 0.0, yleft = geo->yhop_l, yright = geo->yhop_r;
  double zbottom = geo->hop_height;
  for (int64_t i = 0; i < nn; i++) {
    const avf_particle *p = &pp[i];
    avf_double3 ff = force[i], tq = torque[i];
    double ri = p->radius, rho = p->density;
    double mass = ORC_SPH_FACT * ri * ri * ri * rho;
    double km = 0.5 * kappa * mass, gm = 0.5 * gamma * mass;
    for (int ns = 0; ns < 4; ns++) {
      avf_double3 r = orc_sphere(p, ns), rrc = {0, 0, 0}, nvec = {0, 0, 0};
      double dist, comp = 0.0;
      int touch = 0;
      if (r.z < zbottom || r.x > geo->hop_width) continue;
      switch (nw) {
        case 0:
          dist = r.x - xrear;
          if (dist <= ri) {
            touch = 1;
            rrc.x = xrear; rrc.y = r.y; rrc.z = r.z;
            nvec.x = -1.0; nvec.y = 0.0; nvec.z = 0.0;
            comp = ri - dist;
          }
          dist = xfront - r.x;
          if (dist <= ri) {
            touch = 2;
            rrc.x = xfront; rrc.y = r.y; rrc.z = r.z;
            nvec.x = 1.0; nvec.y = 0.0; nvec.z = 0.0;
            comp = ri - dist;
          }
          break;
        case 1:
          dist = r.y - yleft;
          if (dist <= ri) {
            touch = 3;
            rrc.x = r.x; rrc.y = yleft; rrc.z = r.z;
            nvec.x = nvec.z = 0.0; nvec.y = -1.0;
            comp = ri - dist;
          }
          dist = yright - r.y;
          if (dist <= ri) {
            touch = 4;
            rrc.x = r.x; rrc.y = yright; rrc.z = r.z;
            nvec.x = nvec.z = 0.0; nvec.y = 1.0;
            comp = ri - dist;
          }
          break;
        case 2:
          dist = r.z - zbottom;
          if (dist <= ri) {
            touch = (time < 0.0) ? 5 : 0;
            rrc.x = r.x; rrc.y = r.y; rrc.z = zbottom;
            nvec.x = nvec.y = 0.0; nvec.z = -1.0;
            comp = ri - dist;
          }
          break;
      }
      if (touch) orc_wall_contact(p, rrc, nvec, comp, km, gm, mu, cns_1, dt, mass, 0, &ff, &tq);
    }
    force[i] = ff;
    torque[i] = tq;
  }
}

/* Slope frame -> lab (force_funnel.cu:88-93). */
static inline avf_double3 orc_to_lab(avf_double3 v, double cost, double sint) {
  avf_double3 r;
  r.x = v.x * cost + v.z * sint;
  r.y = v.y;
  r.z = v.z * cost - v.x * sint;
  return r;
}

/* force_funnel.cu:9-213: converging funnel side planes in the slope frame. */
void orc_force_funnel(avf_double3 *force, avf_double3 *torque, const avf_system *sys,
                      const avf_particle *pp, const avf_funnel *fun, const avf_geometry *geo) {
  int64_t nn = sys->npart;
  double cost = sys->cost, sint = sys->sint, mu = sys->mu_pw, kappa = sys->kappa_pw;
  double gamma = sys->gamma_pw, dt = sys->dt, cns_1 = sys->fr_fact_pw;
  double sinfi = fun->sinfi, cosfi = fun->cosfi, tanfi = fun->tanfi;
  for (int64_t i = 0; i < nn; i++) {
    const avf_particle *p = &pp[i];
    avf_double3 ff = force[i], tq = torque[i];
    double ri = p->radius, rho = p->density;
    double mass = ORC_SPH_FACT * ri * ri * ri * rho;
    double km = 0.5 * kappa * mass, gm = 0.5 * gamma * mass;
    for (int ns = 0; ns < 4; ns++) {
      avf_double3 q = orc_sphere_rot(p, ns), rrotc, nvecr, rrc = {0, 0, 0}, nvec = {0, 0, 0};
      double dist, comp = 0.0;
      int touch = 0;
      if (q.x > geo->lfunnel - ri * sinfi) continue;
      dist = fabs(q.y - q.x * tanfi) * cosfi;
      if (dist <= ri) {
        touch = 1;
        rrotc.x = q.x + dist * sinfi;
        rrotc.y = q.y - dist * cosfi;
        rrotc.z = q.z;
        nvecr.x = sinfi;
        nvecr.y = -cosfi;
        nvecr.z = 0.0;
        rrc = orc_to_lab(rrotc, cost, sint);
        nvec = orc_to_lab(nvecr, cost, sint);
        comp = ri - dist;
      }
      dist = fabs(geo->yfunnel - q.y - q.x * tanfi) * cosfi;
      if (dist <= ri) {
        touch = 2;
        rrotc.x = q.x + dist * sinfi;
        rrotc.y = q.y + dist * cosfi;
        rrotc.z = q.z;
        nvecr.x = sinfi;
        nvecr.y = cosfi;
        nvecr.z = 0.0;
        rrc = orc_to_lab(rrotc, cost, sint);
        nvec = orc_to_lab(nvecr, cost, sint);
        comp = ri - dist;
      }
      if (touch) orc_wall_contact(p, rrc, nvec, comp, km, gm, mu, cns_1, dt, mass, 0, &ff, &tq);
    }
    force[i] = ff;
    torque[i] = tq;
  }
}

/* force_funnel.cu:216-404: sphere against the vertical corner edges at the funnel exit. */
void orc_force_fcorners(avf_double3 *force, avf_double3 *torque, const avf_system *sys,
                        const avf_funnel *fun, const avf_particle *pp,
                        const avf_geometry *geo) {
  int64_t nn = sys->npart;
  double mu = sys->mu_pw, kappa = sys->kappa_pw, gamma = sys->gamma_pw, dt = sys->dt;
  double cns_1 = sys->fr_fact_pw, sin_fa = fun->sinfi, yfun_l = sys->dy2;
  double cost = sys->cost, sint = sys->sint;
  double yfun_r = sys->side.y + sys->dy2, hsy = sys->dy2 + sys->side.y / 2.0;
  double lf = geo->lfunnel;
  for (int64_t i = 0; i < nn; i++) {
    const avf_particle *p = &pp[i];
    avf_double3 ff = force[i], tq = torque[i];
    double ri = p->radius, rho = p->density;
    double mass = ORC_SPH_FACT * ri * ri * ri * rho;
    double km = 0.5 * kappa * mass, gm = 0.5 * gamma * mass;
    for (int ns = 0; ns < 4; ns++) {
      avf_double3 q = orc_sphere_rot(p, ns), drr, nvecr, rrotc, rrc, nvec;
      double yc, sep, sep2, dist2, dist, comp;
      if (q.x < lf - ri * sin_fa || q.x > lf + ri) continue;
      if (q.y > yfun_r || q.y < yfun_l) continue;
      yc = (q.y <= hsy) ? yfun_l : yfun_r;
      sep = ri;
      sep2 = sep * sep;
      drr.x = lf - q.x;
      dist2 = drr.x * drr.x;
      if (dist2 > sep2) continue;
      drr.y = yc - q.y;
      dist2 += drr.y * drr.y;
      if (dist2 > sep2) continue;
      dist = sqrt(dist2);
      comp = sep - dist;
      nvecr.x = drr.x / dist;
      nvecr.y = drr.y / dist;
      nvecr.z = 0.0;
      rrotc.x = q.x + ri * nvecr.x;
      rrotc.y = q.y + ri * nvecr.y;
      rrotc.z = q.z;
      rrc = orc_to_lab(rrotc, cost, sint);
      nvec = orc_to_lab(nvecr, cost, sint);
      orc_wall_contact(p, rrc, nvec, comp, km, gm, mu, cns_1, dt, mass, 1, &ff, &tq);
    }
    force[i] = ff;
    torque[i] = tq;
  }
}

/* force_walls.cu:9-199: straight flume side walls for spheres past the funnel
 * (rrot.x > LFUNNEL).  INTER_FLAG = 0 switches dissipation off after the gate opens. */
void orc_force_walls(avf_double3 *force, avf_double3 *torque, const avf_system *sys,
                     const avf_particle *pp, const avf_geometry *geo) {
  int64_t nn = sys->npart;
  double time = sys->time, dt = sys->dt, mu = sys->mu_pw, cns_1 = sys->fr_fact_pw;
  double yleft = sys->dy2, yright = sys->side.y + sys->dy2;
  for (int64_t i = 0; i < nn; i++) {
    const avf_particle *p = &pp[i];
    avf_double3 ff, tq;
    double ri, rho, mass, km, gm;
    if (orc_frozen(time, p->rr.x, sys->side.x)) continue;
    ff = force[i];
    tq = torque[i];
    ri = p->radius;
    rho = p->density;
    mass = ORC_SPH_FACT * ri * ri * ri * rho;
    km = 0.5 * sys->kappa_pw * mass;
    gm = 0.5 * sys->gamma_pw * mass;
    for (int ns = 0; ns < 4; ns++) {
      avf_double3 r = orc_sphere(p, ns), q = orc_sphere_rot(p, ns), rrc = {0, 0, 0};
      avf_double3 nvec = {0, 0, 0};
      double dist, comp = 0.0;
      int touch = 0;
      if (q.x <= geo->lfunnel) continue;
      dist = r.y - yleft;
      if (dist <= ri) {
        touch = 1;
        rrc.x = r.x; rrc.y = yleft; rrc.z = r.z;
        nvec.x = nvec.z = 0.0; nvec.y = -1.0;
        comp = ri - dist;
      }
      dist = yright - r.y;
      if (dist <= ri) {
        touch = 2;
        rrc.x = r.x; rrc.y = yright; rrc.z = r.z;
        nvec.x = nvec.z = 0.0; nvec.y = 1.0;
        comp = ri - dist;
      }
      if (!geo->inter_flag && time > 0.0) gm = mu = cns_1 = 0.0;
      if (touch) orc_wall_contact(p, rrc, nvec, comp, km, gm, mu, cns_1, dt, mass, 0, &ff, &tq);
    }
    force[i] = ff;
    torque[i] = tq;
  }
}

/* force_flume.cu:10-182: inclined floor z = -x tan(theta) (no x extent). */
void orc_force_flume(avf_double3 *force, avf_double3 *torque, const avf_system *sys,
                     const avf_particle *pp) {
  int64_t nn = sys->npart;
  double mu = sys->mu_pw, kappa = sys->kappa_pw, gamma = sys->gamma_pw, dt = sys->dt;
  double sint = sys->sint, cost = sys->cost, cns_1 = sys->fr_fact_pw, tant = sint / cost;
  double time = sys->time;
  for (int64_t i = 0; i < nn; i++) {
    const avf_particle *p = &pp[i];
    avf_double3 ff, tq;
    double ri, rho, mass, km, gm;
    if (orc_frozen(time, p->rr.x, sys->side.x)) continue;
    ff = force[i];
    tq = torque[i];
    ri = p->radius;
    rho = p->density;
    mass = ORC_SPH_FACT * ri * ri * ri * rho;
    km = 0.5 * kappa * mass;
    gm = 0.5 * gamma * mass;
    for (int ns = 0; ns < 4; ns++) {
      avf_double3 r = orc_sphere(p, ns), nvec, rrc;
      double rect = -tant * r.x - r.z, dist = fabs(cost * rect);
      if (!(dist <= ri)) continue;
      nvec.x = -sint;
      nvec.y = 0.0;
      nvec.z = -cost;
      rrc.x = r.x - dist * sint;
      rrc.y = r.y;
      rrc.z = r.z - dist * cost;
      orc_wall_contact(p, rrc, nvec, ri - dist, km, gm, mu, cns_1, dt, mass, 0, &ff, &tq);
    }
    force[i] = ff;
    torque[i] = tq;
  }
}

/* force_dep.cu:10-178 (never called by the reference drivers; xhopper defined by the
 * caller): plane z = tan_d (x - x0 - x1) + z0 with normal (sin_d, 0, -cos_d).  The
 * reference form is dep_sin = sint, dep_cos = cost, x0 = Lflume, x1 = xhopper, z0 = 0. */
void orc_force_dep(avf_double3 *force, avf_double3 *torque, const avf_system *sys,
                   const avf_particle *pp, const avf_geometry *geo) {
  int32_t nn = (int32_t)sys->npart; /* `int nn` in the reference */
  double mu = sys->mu_pw, kappa = sys->kappa_pw, gamma = sys->gamma_pw, dt = sys->dt;
  double sind = geo->dep_sin, cosd = geo->dep_cos, tand = sind / cosd;
  double x0 = geo->dep_x0, x1 = geo->dep_x1, z0 = geo->dep_z0, cns_1 = sys->fr_fact_pw;
  for (int32_t i = 0; i < nn; i++) {
    const avf_particle *p = &pp[i];
    avf_double3 ff = force[i], tq = torque[i];
    double ri = p->radius, rho = p->density;
    double mass = ORC_SPH_FACT * ri * ri * ri * rho;
    double km = 0.5 * kappa * mass, gm = 0.5 * gamma * mass;
    for (int ns = 0; ns < 4; ns++) {
      avf_double3 r = orc_sphere(p, ns), nvec, rrc;
      double rect = tand * (r.x - x0 - x1) - (r.z - z0), dist = cosd * (-rect);
      if (!(dist <= ri)) continue;
      nvec.x = sind;
      nvec.y = 0.0;
      nvec.z = -cosd;
      rrc.x = r.x + dist * sind;
      rrc.y = r.y;
      rrc.z = tand * (rrc.x - x0 - x1) + z0;
      orc_wall_contact(p, rrc, nvec, ri - dist, km, gm, mu, cns_1, dt, mass, 0, &ff, &tq);
    }
    force[i] = ff;
    torque[i] = tq;
  }
}

/* ------------------------------------------------------------------ rigid body ----- */

/* get_accel.cu:9-78: a = F/M - drag |v| v - g z; Euler's equations in the body frame. */
void orc_get_accel(const avf_double3 *force, const avf_double3 *torque, avf_double3 *accel,
                   avf_double3 *ang_acc, const avf_system *sys, const avf_particle *pp,
                   const avf_geometry *geo) {
  int64_t nn = sys->npart;
  for (int64_t i = 0; i < nn; i++) {
    const avf_particle *p = &pp[i];
    double mass, rad, Cxy, Cyz, Czx, adrag, vmag;
    avf_double3 I, vv, w, Rx, Ry, Rz, ff, tql, aa, wb, tb, ab, al;
    if (orc_frozen(sys->time, p->rr.x, sys->side.x)) continue;
    mass = p->mass;
    rad = p->radius;
    I = p->moment_of_inertia;
    vv = p->vv;
    w = p->ww;
    Rx = p->Rmatx;
    Ry = p->Rmaty;
    Rz = p->Rmatz;
    Cxy = (I.x - I.y) / I.z;
    Cyz = (I.y - I.z) / I.x;
    Czx = (I.z - I.x) / I.y;
    ff = force[i];
    tql = torque[i];
    adrag = geo->drag * rad * rad / mass;
    vmag = sqrt(vv.x * vv.x + vv.y * vv.y + vv.z * vv.z);
    adrag *= vmag;
    aa.x = ff.x / mass - adrag * vv.x;
    aa.y = ff.y / mass - adrag * vv.y;
    aa.z = ff.z / mass - adrag * vv.z - geo->grav;
    wb.x = w.x * Rx.x + w.y * Ry.x + w.z * Rz.x;
    wb.y = w.x * Rx.y + w.y * Ry.y + w.z * Rz.y;
    wb.z = w.x * Rx.z + w.y * Ry.z + w.z * Rz.z;
    tb.x = tql.x * Rx.x + tql.y * Ry.x + tql.z * Rz.x;
    tb.y = tql.x * Rx.y + tql.y * Ry.y + tql.z * Rz.y;
    tb.z = tql.x * Rx.z + tql.y * Ry.z + tql.z * Rz.z;
    ab.x = tb.x / I.x + Cyz * wb.y * wb.z;
    ab.y = tb.y / I.y + Czx * wb.z * wb.x;
    ab.z = tb.z / I.z + Cxy * wb.x * wb.y;
    al.x = ab.x * Rx.x + ab.y * Rx.y + ab.z * Rx.z;
    al.y = ab.x * Ry.x + ab.y * Ry.y + ab.z * Ry.z;
    al.z = ab.x * Rz.x + ab.y * Rz.y + ab.z * Rz.z;
    accel[i] = aa;
    ang_acc[i] = al;
  }
}

/* init_Verlet.cu:11-152: position update, velocity predict, Rodrigues rotation of R,
 * per-column renormalisation, angular velocity predict, sphere centres.  A clump past
 * side.x (time > 0) is frozen: radius = 0, nothing else changes. */
void orc_init_Verlet(const avf_double3 *accel, const avf_double3 *ang_acc,
                     const avf_system *sys, avf_particle *pp, const avf_geometry *geo) {
  int64_t nn = sys->npart;
  double dt = sys->dt, hdt = sys->hdt, time = sys->time;
  for (int64_t i = 0; i < nn; i++) {
    avf_particle *p = &pp[i];
    avf_double3 rr = p->rr, vv, aa, Rx, Ry, Rz, ww, anga, dO, n, dRx, dRy, dRz, Tx, Ty, Tz;
    double mag, cdO, sdO, omcdO, norm, rad, pos;
    if (orc_frozen(time, rr.x, sys->side.x)) {
      p->radius = 0.0;
      continue;
    }
    vv = p->vv;
    aa = accel[i];
    Rx = p->Rmatx;
    Ry = p->Rmaty;
    Rz = p->Rmatz;
    ww = p->ww;
    rad = p->radius;
    anga = ang_acc[i];

    rr.x += dt * (vv.x + hdt * aa.x);
    rr.y += dt * (vv.y + hdt * aa.y);
    rr.z += dt * (vv.z + hdt * aa.z);
    vv.x += dt * aa.x;
    vv.y += dt * aa.y;
    vv.z += dt * aa.z;

    dO.x = dt * (ww.x + hdt * anga.x);
    dO.y = dt * (ww.y + hdt * anga.y);
    dO.z = dt * (ww.z + hdt * anga.z);
    mag = sqrt(dO.x * dO.x + dO.y * dO.y + dO.z * dO.z);
    if (mag > 0.0) {
      n.x = dO.x / mag;
      n.y = dO.y / mag;
      n.z = dO.z / mag;
      cdO = cos(mag);
      sdO = sin(mag);
      omcdO = 1.0 - cdO;
      dRx.x = omcdO * n.x * n.x + cdO;
      dRx.y = omcdO * n.x * n.y - sdO * n.z;
      dRx.z = omcdO * n.x * n.z + sdO * n.y;
      dRy.x = omcdO * n.y * n.x + sdO * n.z;
      dRy.y = omcdO * n.y * n.y + cdO;
      dRy.z = omcdO * n.y * n.z - sdO * n.x;
      dRz.x = omcdO * n.z * n.x - sdO * n.y;
      dRz.y = omcdO * n.z * n.y + sdO * n.x;
      dRz.z = omcdO * n.z * n.z + cdO;
    } else {
      dRx.x = 1.0; dRx.y = 0.0; dRx.z = 0.0;
      dRy.x = 0.0; dRy.y = 1.0; dRy.z = 0.0;
      dRz.x = 0.0; dRz.y = 0.0; dRz.z = 1.0;
    }
    Tx.x = dRx.x * Rx.x + dRx.y * Ry.x + dRx.z * Rz.x;
    Tx.y = dRx.x * Rx.y + dRx.y * Ry.y + dRx.z * Rz.y;
    Tx.z = dRx.x * Rx.z + dRx.y * Ry.z + dRx.z * Rz.z;
    Ty.x = dRy.x * Rx.x + dRy.y * Ry.x + dRy.z * Rz.x;
    Ty.y = dRy.x * Rx.y + dRy.y * Ry.y + dRy.z * Rz.y;
    Ty.z = dRy.x * Rx.z + dRy.y * Ry.z + dRy.z * Rz.z;
    Tz.x = dRz.x * Rx.x + dRz.y * Ry.x + dRz.z * Rz.x;
    Tz.y = dRz.x * Rx.y + dRz.y * Ry.y + dRz.z * Rz.y;
    Tz.z = dRz.x * Rx.z + dRz.y * Ry.z + dRz.z * Rz.z;

    norm = sqrt(Tx.x * Tx.x + Ty.x * Ty.x + Tz.x * Tz.x);
    Rx.x = Tx.x / norm;
    Ry.x = Ty.x / norm;
    Rz.x = Tz.x / norm;
    norm = sqrt(Tx.y * Tx.y + Ty.y * Ty.y + Tz.y * Tz.y);
    Rx.y = Tx.y / norm;
    Ry.y = Ty.y / norm;
    Rz.y = Tz.y / norm;
    norm = sqrt(Tx.z * Tx.z + Ty.z * Ty.z + Tz.z * Tz.z);
    Rx.z = Tx.z / norm;
    Ry.z = Ty.z / norm;
    Rz.z = Tz.z / norm;

    ww.x += dt * anga.x;
    ww.y += dt * anga.y;
    ww.z += dt * anga.z;

    p->rr = rr;
    p->vv = vv;
    p->Rmatx = Rx;
    p->Rmaty = Ry;
    p->Rmatz = Rz;
    p->ww = ww;

    pos = rad / geo->sphere_div;
    p->rr_a.x = rr.x + pos * (Rx.x + Rx.y + Rx.z);
    p->rr_a.y = rr.y + pos * (Ry.x + Ry.y + Ry.z);
    p->rr_a.z = rr.z + pos * (Rz.x + Rz.y + Rz.z);
    p->rr_b.x = rr.x + pos * (Rx.x - Rx.y - Rx.z);
    p->rr_b.y = rr.y + pos * (Ry.x - Ry.y - Ry.z);
    p->rr_b.z = rr.z + pos * (Rz.x - Rz.y - Rz.z);
    p->rr_c.x = rr.x + pos * (-Rx.x - Rx.y + Rx.z);
    p->rr_c.y = rr.y + pos * (-Ry.x - Ry.y + Ry.z);
    p->rr_c.z = rr.z + pos * (-Rz.x - Rz.y + Rz.z);
    p->rr_d.x = rr.x + pos * (-Rx.x + Rx.y - Rx.z);
    p->rr_d.y = rr.y + pos * (-Ry.x + Ry.y - Ry.z);
    p->rr_d.z = rr.z + pos * (-Rz.x + Rz.y - Rz.z);
  }
}

/* rotate_vectors.cu:7-56: lab -> slope frame for the centre and the four spheres. */
void orc_rotate_vectors(avf_particle *pp, const avf_system *sys) {
  int64_t nn = sys->npart;
  double sint = sys->sint, cost = sys->cost;
  for (int64_t i = 0; i < nn; i++) {
    avf_particle *p = &pp[i];
    avf_double3 in[5] = {p->rr, p->rr_a, p->rr_b, p->rr_c, p->rr_d}, out[5];
    for (int k = 0; k < 5; k++) {
      out[k].x = in[k].x * cost - in[k].z * sint;
      out[k].y = in[k].y;
      out[k].z = in[k].x * sint + in[k].z * cost;
    }
    p->rrot = out[0];
    p->rrot_a = out[1];
    p->rrot_b = out[2];
    p->rrot_c = out[3];
    p->rrot_d = out[4];
  }
}

/* finish_Verlet.cu:10-56: velocity corrections; a frozen clump gets vv = 0. */
void orc_finish_Verlet(const avf_double3 *accel_i, const avf_double3 *ang_acc_i,
                       const avf_double3 *accel_f, const avf_double3 *ang_acc_f,
                       const avf_system *sys, avf_particle *pp) {
  int64_t nn = sys->npart;
  double hdt = sys->hdt;
  for (int64_t i = 0; i < nn; i++) {
    avf_particle *p = &pp[i];
    avf_double3 vv, ww, ai, af, gi, gf;
    if (orc_frozen(sys->time, p->rr.x, sys->side.x)) {
      p->vv.x = p->vv.y = p->vv.z = 0.0;
      continue;
    }
    vv = p->vv;
    ww = p->ww;
    ai = accel_i[i];
    gi = ang_acc_i[i];
    af = accel_f[i];
    gf = ang_acc_f[i];
    vv.x += hdt * (af.x - ai.x);
    vv.y += hdt * (af.y - ai.y);
    vv.z += hdt * (af.z - ai.z);
    ww.x += hdt * (gf.x - gi.x);
    ww.y += hdt * (gf.y - gi.y);
    ww.z += hdt * (gf.z - gi.z);
    p->vv = vv;
    p->ww = ww;
  }
}

/* ------------------------------------------------------------------ step driver ---- */

int orc_sim_init(orc_sim *s, const avf_system *sys, const avf_geometry *geo,
                 const avf_funnel *fun, int64_t ntags, int32_t phase, const avf_particle *pp) {
  int64_t n = sys->npart;
  memset(s, 0, sizeof(*s));
  s->sys = *sys;
  s->geo = *geo;
  s->fun = *fun;
  s->cell.ntags = ntags;
  s->phase = phase;
  s->n = n;
  s->pp = (avf_particle *)malloc((size_t)n * sizeof(avf_particle));
  s->ff = (avf_double3 *)calloc((size_t)n, sizeof(avf_double3));
  s->tq = (avf_double3 *)calloc((size_t)n, sizeof(avf_double3));
  s->aa_i = (avf_double3 *)calloc((size_t)n, sizeof(avf_double3));
  s->anga_i = (avf_double3 *)calloc((size_t)n, sizeof(avf_double3));
  s->aa_f = (avf_double3 *)calloc((size_t)n, sizeof(avf_double3));
  s->anga_f = (avf_double3 *)calloc((size_t)n, sizeof(avf_double3));
  if (!s->pp || !s->ff || !s->tq || !s->aa_i || !s->anga_i || !s->aa_f || !s->anga_f)
    return AVF_ERR_NOMEM;
  memcpy(s->pp, pp, (size_t)n * sizeof(avf_particle));
  return 0;
}

void orc_sim_free(orc_sim *s) {
  free(s->pp);
  free(s->ff);
  free(s->tq);
  free(s->aa_i);
  free(s->anga_i);
  free(s->aa_f);
  free(s->anga_f);
  free(s->cell_vec);
  free(s->nocup);
  memset(s, 0, sizeof(*s));
}

int orc_sim_regrid(orc_sim *s) {
  orc_cell_setup(&s->cell, s->pp, &s->sys, &s->geo);
  if (s->cell.ncell3 <= 0 || s->cell.ntot <= 0) return AVF_ERR_ARG;
  if (s->cell.ntot > s->cell_cap) {
    free(s->cell_vec);
    s->cell_vec = (int64_t *)malloc((size_t)s->cell.ntot * sizeof(int64_t));
    s->cell_cap = s->cell.ntot;
    if (!s->cell_vec) return AVF_ERR_NOMEM;
  }
  if (s->cell.ncell3 > s->ocup_cap) {
    free(s->nocup);
    s->nocup = (int32_t *)malloc((size_t)s->cell.ncell3 * sizeof(int32_t));
    s->ocup_cap = s->cell.ncell3;
    if (!s->nocup) return AVF_ERR_NOMEM;
  }
  return 0;
}

static void orc_pp_passes(orc_sim *s) {
  for (int sph = 1; sph <= 16; sph++)
    orc_force_part(s->ff, s->tq, s->cell_vec, s->nocup, s->pp, &s->sys, &s->cell, sph, s->rec);
}

/* Fill: avalanche_hst.cu:445-461.  Avalanche: avalanche_hst.cu:678-704 (+ the deposit
 * floor after the flume when enabled). */
/* One force stage.  initial != 0: the initial force pass of avalanche_hst.cu:365-384, which
 * runs force_part and then force_hop in either phase (the filling composition). */
static int orc_sim_forces_pass(orc_sim *s, int initial) {
  int rc;
  int32_t mods = s->phase == AVF_PHASE_FILL ? s->geo.modules_fill : s->geo.modules_avalanche;
  const int fill = s->phase == AVF_PHASE_FILL || initial;
  if (initial) mods &= AVF_MOD_PP | AVF_MOD_HOP;
  orc_clear_forces(s->n, s->ff, s->tq);
  orc_clear_cell(&s->cell, s->cell_vec);
  orc_clear_ocup(&s->cell, s->nocup);
  rc = orc_cell_locate(&s->cell, &s->sys, s->pp, s->nocup, s->cell_vec, &s->last_bad);
  if (rc) {
    s->last_error = rc;
    return rc;
  }
  if (s->rec) s->rec->count = 0;
  if (fill) {
    if (mods & AVF_MOD_PP) orc_pp_passes(s);
    if (mods & AVF_MOD_HOP)
      for (int nw = 0; nw < 3; nw++) orc_force_hop(s->ff, s->tq, &s->sys, s->pp, &s->geo, nw);
  } else {
    if (mods & AVF_MOD_HOP)
      for (int nw = 0; nw < 3; nw++) orc_force_hop(s->ff, s->tq, &s->sys, s->pp, &s->geo, nw);
    if (mods & AVF_MOD_FUNNEL) orc_force_funnel(s->ff, s->tq, &s->sys, s->pp, &s->fun, &s->geo);
    if (mods & AVF_MOD_FCORNERS)
      orc_force_fcorners(s->ff, s->tq, &s->sys, &s->fun, s->pp, &s->geo);
    if (mods & AVF_MOD_WALLS) orc_force_walls(s->ff, s->tq, &s->sys, s->pp, &s->geo);
    if (mods & AVF_MOD_PP) orc_pp_passes(s);
    if (mods & AVF_MOD_FLUME) orc_force_flume(s->ff, s->tq, &s->sys, s->pp);
    if (mods & AVF_MOD_DEP) orc_force_dep(s->ff, s->tq, &s->sys, s->pp, &s->geo);
  }
  return 0;
}

int orc_sim_forces(orc_sim *s) { return orc_sim_forces_pass(s, 0); }

int orc_sim_initial(orc_sim *s) {
  int rc = orc_sim_forces_pass(s, 1);
  if (rc) return rc;
  orc_get_accel(s->ff, s->tq, s->aa_i, s->anga_i, &s->sys, s->pp, &s->geo);
  return 0;
}

int orc_sim_step(orc_sim *s) {
  int rc;
  avf_double3 *sw;
  orc_init_Verlet(s->aa_i, s->anga_i, &s->sys, s->pp, &s->geo);
  s->sys.time += s->sys.dt;
  if (s->phase == AVF_PHASE_AVALANCHE) orc_rotate_vectors(s->pp, &s->sys);
  rc = orc_sim_forces(s);
  if (rc) return rc;
  orc_get_accel(s->ff, s->tq, s->aa_f, s->anga_f, &s->sys, s->pp, &s->geo);
  orc_finish_Verlet(s->aa_i, s->anga_i, s->aa_f, s->anga_f, &s->sys, s->pp);
  if (s->geo.regrid_interval > 0 && s->cc % s->geo.regrid_interval == 0) {
    rc = orc_sim_regrid(s);
    if (rc) return rc;
  }
  sw = s->aa_i; s->aa_i = s->aa_f; s->aa_f = sw;
  sw = s->anga_i; s->anga_i = s->anga_f; s->anga_f = sw;
  s->cc++;
  return 0;
}

/* ------------------------------------------------------------------ brute force ---- */

static int orc_contact_cmp(const void *a, const void *b) {
  const avf_contact *x = (const avf_contact *)a, *y = (const avf_contact *)b;
  if (x->l != y->l) return x->l < y->l ? -1 : 1;
  if (x->sm != y->sm) return x->sm < y->sm ? -1 : 1;
  if (x->sl != y->sl) return x->sl < y->sl ? -1 : 1;
  if (x->m != y->m) return x->m < y->m ? -1 : 1;
  return 0;
}

int64_t orc_brute_contacts(const avf_particle *pp, int64_t n, const avf_system *sys,
                           avf_contact *buf, int64_t cap) {
  int64_t count = 0;
  for (int64_t l = 0; l < n; l++) {
    if (orc_frozen(sys->time, pp[l].rr.x, sys->side.x)) continue;
    for (int64_t m = 0; m < n; m++) {
      if (m == l || orc_frozen(sys->time, pp[m].rr.x, sys->side.x)) continue;
      for (int sm = 0; sm < 4; sm++)
        for (int sl = 0; sl < 4; sl++) {
          avf_double3 r1 = orc_sphere(&pp[l], sl), r2 = orc_sphere(&pp[m], sm);
          double sep = pp[l].radius + pp[m].radius, sep2 = sep * sep;
          double dx = r2.x - r1.x, dy = r2.y - r1.y, dz = r2.z - r1.z;
          double d2 = dx * dx + dy * dy + dz * dz;
          if (d2 < sep2) {
            if (count < cap) {
              buf[count].l = l;
              buf[count].m = m;
              buf[count].sl = sl;
              buf[count].sm = sm;
              buf[count].repulsive = -1; /* not evaluated here */
              buf[count].reserved = 0;
            }
            count++;
          }
        }
    }
  }
  if (buf && count <= cap) qsort(buf, (size_t)count, sizeof(avf_contact), orc_contact_cmp);
  return count;
}

/* ------------------------------------------------------------- handle helpers ---- */

orc_sim *orc_sim_new(const avf_system *sys, const avf_geometry *geo, const avf_funnel *fun,
                     int64_t ntags, int32_t phase, const avf_particle *pp) {
  orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
  if (!s) return NULL;
  if (orc_sim_init(s, sys, geo, fun, ntags, phase, pp) || orc_sim_regrid(s)) {
    orc_sim_free(s);
    free(s);
    return NULL;
  }
  return s;
}

void orc_sim_delete(orc_sim *s) {
  if (!s) return;
  free(s->rec);
  orc_sim_free(s);
  free(s);
}

int orc_sim_steps(orc_sim *s, int64_t nsteps) {
  for (int64_t k = 0; k < nsteps; k++) {
    int rc = orc_sim_step(s);
    if (rc) return rc;
  }
  return 0;
}

/* Hopper opening, avalanche_hst.cu:606-613. */
void orc_sim_open(orc_sim *s, double xfinal) {
  s->sys.side.x = xfinal;
  s->phase = AVF_PHASE_AVALANCHE;
  s->cc = 0;
  s->sys.time = 0.0;
}

void orc_sim_get(const orc_sim *s, avf_particle *pp, avf_double3 *aa_i, avf_double3 *anga_i,
                 avf_double3 *ff, avf_double3 *tq, avf_system *sys, avf_cell *cell) {
  size_t n = (size_t)s->n;
  if (pp) memcpy(pp, s->pp, n * sizeof(avf_particle));
  if (aa_i) memcpy(aa_i, s->aa_i, n * sizeof(avf_double3));
  if (anga_i) memcpy(anga_i, s->anga_i, n * sizeof(avf_double3));
  if (ff) memcpy(ff, s->ff, n * sizeof(avf_double3));
  if (tq) memcpy(tq, s->tq, n * sizeof(avf_double3));
  if (sys) *sys = s->sys;
  if (cell) *cell = s->cell;
}

void orc_sim_set(orc_sim *s, const avf_particle *pp, const avf_double3 *aa_i,
                 const avf_double3 *anga_i) {
  size_t n = (size_t)s->n;
  if (pp) memcpy(s->pp, pp, n * sizeof(avf_particle));
  if (aa_i) memcpy(s->aa_i, aa_i, n * sizeof(avf_double3));
  if (anga_i) memcpy(s->anga_i, anga_i, n * sizeof(avf_double3));
}

/* Records the directed contacts of every later pp stage (count reset per stage). */
int orc_sim_record(orc_sim *s, avf_contact *buf, int64_t cap) {
  if (!s->rec) s->rec = (orc_contacts *)calloc(1, sizeof(orc_contacts));
  if (!s->rec) return AVF_ERR_NOMEM;
  s->rec->buf = buf;
  s->rec->cap = cap;
  s->rec->count = 0;
  return 0;
}

int64_t orc_sim_recorded(const orc_sim *s) { return s->rec ? s->rec->count : 0; }
int32_t orc_sim_error(const orc_sim *s, int64_t *bad) {
  if (bad) *bad = s->last_bad;
  return s->last_error;
}
