/*
 * ref_entry.cpp -- plain-pointer entry points into the reference's own host ops
 * (TEST INFRASTRUCTURE ONLY).
 *
 * Linked by oracle/Makefile against the reference's *_hst sources, compiled unchanged from
 * /root/reference with hipcc (which supplies double3 / long3 / __global__ the way nvcc
 * does).  Output goes to oracle/_ref/libref_hst.so, which tests use to pin the oracle.
 * The step composition below follows avalanche_hst.cu:365-384 (initial forces),
 * :438-467 + :589-595 (hopper filling) and :668-710 + :844-850 (avalanche).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "cabecera.hu"

static_assert(sizeof(particle) == 408, "particle layout");
static_assert(sizeof(system_par) == 256, "system_par layout");
static_assert(sizeof(cell_par) == 168, "cell_par layout");
static_assert(sizeof(Funnel) == 24, "Funnel layout");
static_assert(sizeof(double3) == 24, "double3 layout");

#define S(p) (*(const system_par *)(p))
#define C(p) (*(const cell_par *)(p))
#define F(p) (*(const Funnel *)(p))
#define D3(p) ((double3 *)(p))
#define PP(p) ((particle *)(p))

extern "C" {

int ref_sizes(int64_t *out) {
  out[0] = sizeof(particle);
  out[1] = sizeof(system_par);
  out[2] = sizeof(cell_par);
  out[3] = sizeof(Funnel);
  return 0;
}

void ref_clear_forces(const void *sys, void *f, void *t) { clear_forces_hst(S(sys), D3(f), D3(t)); }
void ref_clear_cell(const void *cell, long *cell_vec) { clear_cell_hst(C(cell), cell_vec); }
void ref_clear_ocup(const void *cell, int *nocup) { clear_ocup_hst(C(cell), nocup); }
void ref_cell_locate(const void *cell, const void *sys, void *pp, int *nocup, long *cell_vec) {
  cell_locate_hst(C(cell), S(sys), PP(pp), nocup, cell_vec);
}
void ref_cell_dim(void *rr_max, void *rr_min, void *pp, const void *sys) {
  cell_dim(D3(rr_max), D3(rr_min), PP(pp), S(sys));
}
void ref_force_part(void *f, void *t, long *cell_vec, int *nocup, void *pp, const void *sys,
                    const void *cell, int sph_id) {
  force_part_hst(D3(f), D3(t), cell_vec, nocup, PP(pp), S(sys), C(cell), sph_id);
}
void ref_force_hop(void *f, void *t, const void *sys, void *pp, int nw) {
  force_hop_hst(D3(f), D3(t), S(sys), PP(pp), (char)nw);
}
void ref_force_funnel(void *f, void *t, const void *sys, void *pp, const void *fun) {
  force_funnel_hst(D3(f), D3(t), S(sys), PP(pp), F(fun));
}
void ref_force_fcorners(void *f, void *t, const void *sys, const void *fun, void *pp) {
  force_fcorners_hst(D3(f), D3(t), S(sys), F(fun), PP(pp));
}
void ref_force_walls(void *f, void *t, const void *sys, void *pp) {
  force_walls_hst(D3(f), D3(t), S(sys), PP(pp));
}
void ref_force_flume(void *f, void *t, const void *sys, void *pp) {
  force_flume_hst(D3(f), D3(t), S(sys), PP(pp));
}
void ref_force_dep(void *f, void *t, const void *sys, void *pp) {
  force_dep_hst(D3(f), D3(t), S(sys), PP(pp));
}
void ref_get_accel(void *f, void *t, void *a, void *al, const void *sys, void *pp) {
  get_accel_hst(D3(f), D3(t), D3(a), D3(al), S(sys), PP(pp));
}
void ref_init_Verlet(void *a, void *al, const void *sys, void *pp) {
  init_Verlet_hst(D3(a), D3(al), S(sys), PP(pp));
}
void ref_rotate_vectors(void *pp, const void *sys) { rotate_vectors_hst(PP(pp), S(sys)); }
void ref_finish_Verlet(void *ai, void *gi, void *af, void *gf, const void *sys, void *pp) {
  finish_Verlet_hst(D3(ai), D3(gi), D3(af), D3(gf), S(sys), PP(pp));
}
void ref_generate_particles(void *pp, const void *sys) { generate_particles(PP(pp), S(sys)); }
void ref_calculate_particle_parameters(const void *sys, void *pp) {
  calculate_particle_parameters_hst(S(sys), PP(pp));
}
void ref_place_particles(const void *sys, void *pp) { place_particles(S(sys), PP(pp)); }
void ref_set_initial_velocity(const void *sys, void *pp) { set_initial_velocity(S(sys), PP(pp)); }
double ref_rannew64(long *idum) { return rannew64(idum); }
int ref_error_check(const void *sys, char *msg) { return error_check(S(sys), msg); }

/* ------------------------------------------------------------- step composition ---- */

typedef struct ref_sim {
  system_par sys;
  Funnel fun;
  cell_par cell;
  long n;
  int phase; /* 0 = filling, 1 = avalanche */
  long cc;
  particle *pp;
  double3 *ff, *tq, *aa_i, *anga_i, *aa_f, *anga_f;
  long *cell_vec;
  int *nocup;
  long cell_cap, ocup_cap;
} ref_sim;

/* avalanche_hst.cu:115 */
static const double kDist = GYR + 0.05f;

/* avalanche_hst.cu:311-333 */
static void ref_regrid(ref_sim *s) {
  cell_par *c = &s->cell;
  cell_dim(&c->rr_max, &c->rr_min, s->pp, s->sys);
  c->length.x = c->rr_max.x - c->rr_min.x;
  c->length.y = c->rr_max.y - c->rr_min.y;
  c->length.z = c->rr_max.z - c->rr_min.z;
  c->ncell.x = (long)(c->length.x / (kDist * s->sys.rmax));
  c->ncell.y = (long)(c->length.y / (kDist * s->sys.rmax));
  c->ncell.z = (long)(c->length.z / (kDist * s->sys.rmax));
  c->cell_side.x = c->length.x / ((double)c->ncell.x);
  c->cell_side.y = c->length.y / ((double)c->ncell.y);
  c->cell_side.z = c->length.z / ((double)c->ncell.z);
  c->cs_inv.x = 1.0 / c->cell_side.x;
  c->cs_inv.y = 1.0 / c->cell_side.y;
  c->cs_inv.z = 1.0 / c->cell_side.z;
  c->ncell3 = c->ncell.x * c->ncell.y * c->ncell.z;
  c->ntot = c->ncell3 * c->ntags;
  if (c->ntot > s->cell_cap) {
    free(s->cell_vec);
    s->cell_vec = (long *)malloc(c->ntot * sizeof(long));
    s->cell_cap = c->ntot;
  }
  if (c->ncell3 > s->ocup_cap) {
    free(s->nocup);
    s->nocup = (int *)malloc(c->ncell3 * sizeof(int));
    s->ocup_cap = c->ncell3;
  }
}

void *ref_sim_create(const void *sys, const void *fun, long ntags, int phase, const void *pp) {
  ref_sim *s = (ref_sim *)calloc(1, sizeof(ref_sim));
  s->sys = S(sys);
  s->fun = F(fun);
  s->cell.ntags = ntags;
  s->phase = phase;
  s->n = s->sys.npart;
  s->pp = (particle *)malloc(s->n * sizeof(particle));
  memcpy(s->pp, pp, s->n * sizeof(particle));
  s->ff = (double3 *)calloc(s->n, sizeof(double3));
  s->tq = (double3 *)calloc(s->n, sizeof(double3));
  s->aa_i = (double3 *)calloc(s->n, sizeof(double3));
  s->anga_i = (double3 *)calloc(s->n, sizeof(double3));
  s->aa_f = (double3 *)calloc(s->n, sizeof(double3));
  s->anga_f = (double3 *)calloc(s->n, sizeof(double3));
  ref_regrid(s);
  return s;
}

void ref_sim_destroy(void *h) {
  ref_sim *s = (ref_sim *)h;
  free(s->pp); free(s->ff); free(s->tq); free(s->aa_i); free(s->anga_i); free(s->aa_f);
  free(s->anga_f); free(s->cell_vec); free(s->nocup); free(s);
}

/* initial != 0: the initial force pass of avalanche_hst.cu:365-384 (force_part, then
 * force_hop, in either phase) */
static void ref_forces(ref_sim *s, int initial = 0) {
  clear_forces_hst(s->sys, s->ff, s->tq);
  clear_cell_hst(s->cell, s->cell_vec);
  clear_ocup_hst(s->cell, s->nocup);
  cell_locate_hst(s->cell, s->sys, s->pp, s->nocup, s->cell_vec);
  if (s->phase == 0 || initial) {
    for (int i = 1; i <= 16; i++)
      force_part_hst(s->ff, s->tq, s->cell_vec, s->nocup, s->pp, s->sys, s->cell, i);
    for (int i = 0; i < 3; i++) force_hop_hst(s->ff, s->tq, s->sys, s->pp, (char)i);
  } else {
    for (int i = 0; i < 3; i++) force_hop_hst(s->ff, s->tq, s->sys, s->pp, (char)i);
    force_funnel_hst(s->ff, s->tq, s->sys, s->pp, s->fun);
    force_fcorners_hst(s->ff, s->tq, s->sys, s->fun, s->pp);
    force_walls_hst(s->ff, s->tq, s->sys, s->pp);
    for (int i = 1; i <= 16; i++)
      force_part_hst(s->ff, s->tq, s->cell_vec, s->nocup, s->pp, s->sys, s->cell, i);
    force_flume_hst(s->ff, s->tq, s->sys, s->pp);
  }
}

/* avalanche_hst.cu:365-384 */
void ref_sim_initial(void *h) {
  ref_sim *s = (ref_sim *)h;
  ref_forces(s, 1);
  get_accel_hst(s->ff, s->tq, s->aa_i, s->anga_i, s->sys, s->pp);
}

void ref_sim_step(void *h, long nsteps) {
  ref_sim *s = (ref_sim *)h;
  for (long k = 0; k < nsteps; k++) {
    init_Verlet_hst(s->aa_i, s->anga_i, s->sys, s->pp);
    s->sys.time += s->sys.dt;
    if (s->phase == 1) rotate_vectors_hst(s->pp, s->sys);
    ref_forces(s);
    get_accel_hst(s->ff, s->tq, s->aa_f, s->anga_f, s->sys, s->pp);
    finish_Verlet_hst(s->aa_i, s->anga_i, s->aa_f, s->anga_f, s->sys, s->pp);
    if (s->cc % 2000 == 0) ref_regrid(s);
    double3 *sw = s->aa_i; s->aa_i = s->aa_f; s->aa_f = sw;
    sw = s->anga_i; s->anga_i = s->anga_f; s->anga_f = sw;
    s->cc++;
  }
}

/* Hopper opening, avalanche_hst.cu:606-613. */
void ref_sim_open(void *h, double xfinal) {
  ref_sim *s = (ref_sim *)h;
  s->sys.side.x = xfinal;
  s->phase = 1;
  s->cc = 0;
  s->sys.time = 0.0;
}

void ref_sim_get(void *h, void *pp, void *aa_i, void *anga_i, void *ff, void *tq, void *sys,
                 void *cell) {
  ref_sim *s = (ref_sim *)h;
  if (pp) memcpy(pp, s->pp, s->n * sizeof(particle));
  if (aa_i) memcpy(aa_i, s->aa_i, s->n * sizeof(double3));
  if (anga_i) memcpy(anga_i, s->anga_i, s->n * sizeof(double3));
  if (ff) memcpy(ff, s->ff, s->n * sizeof(double3));
  if (tq) memcpy(tq, s->tq, s->n * sizeof(double3));
  if (sys) memcpy(sys, &s->sys, sizeof(system_par));
  if (cell) memcpy(cell, &s->cell, sizeof(cell_par));
}

}  // extern "C"
