/*
 * dem_oracle.h -- CPU oracle for the DEM step (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C restatement of the reference's host (`_hst`) step ops, in the same operation
 * order and with the same literals, generalised only through avf_geometry (whose reference
 * values reproduce the reference exactly).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never links it.
 *
 * Pinning.  The reference's own sources cannot be built here (they need CUDA's runtime
 * header), so every function is pinned by checks that the mathematics or the reference's
 * definitions fix, each chosen so that a plausible slip fails it (tests/oracle_mutations.py
 * plants 22 such slips; every one is caught):
 *   orc_force_hop        hopper walls and gate at rest / approaching, gate only for t < 0,
 *                        walls skipped below the bottom, second wall wins
 *                        (tests/test_oracle_pins.py: test_hopper_*)
 *   orc_force_funnel     plane force at rest (test_closed_forms.py: test_funnel_plane_*)
 *   orc_force_fcorners   radial corner force and contact point (test_funnel_corner_*)
 *   orc_force_walls      side walls, INTER_FLAG on / off, frozen skip (test_side_walls_*)
 *   orc_force_flume      resting clump carries M g (test_closed_forms.py), tangential law
 *                        on a wall: reset / stick / slide (test_tangential_law_on_a_wall)
 *   orc_force_dep        deposit plane at three tilts (test_deposit_plane)
 *   orc_force_part       contact set vs O(N^2) enumeration (test_closed_forms.py), normal
 *                        law with unequal radii and m_eff, contact point and torque, host
 *                        friction rule (test_normal_law_*, test_pp_contact_point_*,
 *                        test_tangential_law_between_spheres_host_rule), kappa / gamma from
 *                        t_coll and eps by contact time and restitution with unequal radii
 *                        (test_pp_contact_time_*), summation order over pairings, cells and
 *                        neighbours (test_pp_summation_order_*)
 *   orc_sim_forces       module order of both phases (test_module_summation_order,
 *                        test_fill_stage_order_pp_then_hop), wall kappa / gamma by contact
 *                        time and restitution (test_wall_contact_time_*)
 *   orc_cell_dim / orc_cell_setup   margins and grid (test_cell_dim_margins_and_grid)
 *   orc_cell_locate      each clump in its geometric cell, ascending, frozen skipped
 *                        (test_cell_locate_places_each_clump_in_its_geometric_cell)
 *   orc_get_accel        free fall, drag (test_closed_forms.py), Euler's equations with an
 *                        anisotropic inertia (test_euler_equations_anisotropic_inertia)
 *   orc_init_Verlet / orc_finish_Verlet   ballistic flight, Rodrigues vs expm
 *                        (test_closed_forms.py), frozen exit (test_frozen_exit), the
 *                        integrator's restitution (test_pp_contact_time_*)
 *   orc_rotate_vectors   slope frame of points on and above the floor
 *                        (test_rotate_vectors_slope_frame)
 *   orc_brute_contacts   itself a checker (plain O(N^2) enumeration)
 */
#ifndef DEM_ORACLE_H
#define DEM_ORACLE_H

#include "../include/avflume.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Literals of cabecera.hu:5-12 (kept as written: ROOT2 is 1 ulp below sqrt(2),
 * SPH_FACT 1 ulp above 4*pi/3). */
#define ORC_PI 3.1415926535897932
#define ORC_GYR 4.4494897427831781
#define ORC_SPH_FACT 4.188790204786391
#define ORC_ROOT2 1.414213562373095

void orc_clear_forces(int64_t n, avf_double3 *force, avf_double3 *torque);
void orc_clear_cell(const avf_cell *cell, int64_t *cell_vec);
void orc_clear_ocup(const avf_cell *cell, int32_t *nocup);
/* Returns 0, or AVF_ERR_CELL_NEG / AVF_ERR_CELL_FULL with the clump index in *bad
 * (the reference prints and exits there). */
int orc_cell_locate(const avf_cell *cell, const avf_system *sys, const avf_particle *pp,
                    int32_t *nocup, int64_t *cell_vec, int64_t *bad);
void orc_cell_dim(avf_double3 *rr_max, avf_double3 *rr_min, const avf_particle *pp,
                  const avf_system *sys, double ddt_steps);
/* cell_dim + the regrid arithmetic of avalanche_hst.cu:311-333. */
void orc_cell_setup(avf_cell *cell, const avf_particle *pp, const avf_system *sys,
                    const avf_geometry *geo);

/* Optional contact recorder for orc_force_part. */
typedef struct orc_contacts {
  avf_contact *buf;
  int64_t cap, count;
} orc_contacts;

void orc_force_part(avf_double3 *force, avf_double3 *torque, const int64_t *cell_vec,
                    const int32_t *nocup, const avf_particle *pp, const avf_system *sys,
                    const avf_cell *cell, int sph_id, orc_contacts *rec);
void orc_force_hop(avf_double3 *force, avf_double3 *torque, const avf_system *sys,
                   const avf_particle *pp, const avf_geometry *geo, int nw);
void orc_force_funnel(avf_double3 *force, avf_double3 *torque, const avf_system *sys,
                      const avf_particle *pp, const avf_funnel *fun, const avf_geometry *geo);
void orc_force_fcorners(avf_double3 *force, avf_double3 *torque, const avf_system *sys,
                        const avf_funnel *fun, const avf_particle *pp,
                        const avf_geometry *geo);
void orc_force_walls(avf_double3 *force, avf_double3 *torque, const avf_system *sys,
                     const avf_particle *pp, const avf_geometry *geo);
void orc_force_flume(avf_double3 *force, avf_double3 *torque, const avf_system *sys,
                     const avf_particle *pp);
void orc_force_dep(avf_double3 *force, avf_double3 *torque, const avf_system *sys,
                   const avf_particle *pp, const avf_geometry *geo);
void orc_get_accel(const avf_double3 *force, const avf_double3 *torque, avf_double3 *accel,
                   avf_double3 *ang_acc, const avf_system *sys, const avf_particle *pp,
                   const avf_geometry *geo);
void orc_init_Verlet(const avf_double3 *accel, const avf_double3 *ang_acc,
                     const avf_system *sys, avf_particle *pp, const avf_geometry *geo);
void orc_rotate_vectors(avf_particle *pp, const avf_system *sys);
void orc_finish_Verlet(const avf_double3 *accel_i, const avf_double3 *ang_acc_i,
                       const avf_double3 *accel_f, const avf_double3 *ang_acc_f,
                       const avf_system *sys, avf_particle *pp);

/* ---------------------------------------------------------------- step driver ------ */

/* Everything one instance of the host loop carries (avalanche_hst.cu:7-28). */
typedef struct orc_sim {
  avf_system sys;
  avf_geometry geo;
  avf_funnel fun;
  avf_cell cell;
  int32_t phase;
  int64_t cc;          /* step counter of the current phase (avalanche_hst.cu:600,852) */
  int64_t n;
  avf_particle *pp;
  avf_double3 *ff, *tq, *aa_i, *anga_i, *aa_f, *anga_f;
  int64_t *cell_vec;
  int32_t *nocup;
  int64_t cell_cap, ocup_cap;
  int32_t last_error;
  int64_t last_bad;
  orc_contacts *rec;   /* optional contact recorder for the pp stage */
} orc_sim;

/* Allocates the arrays of an instance with n clumps (pp copied in). */
int orc_sim_init(orc_sim *s, const avf_system *sys, const avf_geometry *geo,
                 const avf_funnel *fun, int64_t ntags, int32_t phase, const avf_particle *pp);
void orc_sim_free(orc_sim *s);
/* cell_dim + regrid (avalanche_hst.cu:311-333); reallocates cell arrays as needed. */
int orc_sim_regrid(orc_sim *s);
/* The force stage of the current phase on the current state: clear_forces, clear_cell,
 * clear_ocup, cell_locate, then the module order of the phase (fills s->ff, s->tq). */
int orc_sim_forces(orc_sim *s);
/* Initial force pass avalanche_hst.cu:365-384 (forces, then get_accel into aa_i). */
int orc_sim_initial(orc_sim *s);
/* One full step of the current phase, including the cc%regrid_interval regrid and the
 * acceleration swap (avalanche_hst.cu:438-600 / 668-852). */
int orc_sim_step(orc_sim *s);

/* Heap-allocated handle API (used through ctypes by the tests). */
orc_sim *orc_sim_new(const avf_system *sys, const avf_geometry *geo, const avf_funnel *fun,
                     int64_t ntags, int32_t phase, const avf_particle *pp);
void orc_sim_delete(orc_sim *s);
int orc_sim_steps(orc_sim *s, int64_t nsteps);
void orc_sim_open(orc_sim *s, double xfinal);
void orc_sim_get(const orc_sim *s, avf_particle *pp, avf_double3 *aa_i, avf_double3 *anga_i,
                 avf_double3 *ff, avf_double3 *tq, avf_system *sys, avf_cell *cell);
void orc_sim_set(orc_sim *s, const avf_particle *pp, const avf_double3 *aa_i,
                 const avf_double3 *anga_i);
int orc_sim_record(orc_sim *s, avf_contact *buf, int64_t cap);
int64_t orc_sim_recorded(const orc_sim *s);
int32_t orc_sim_error(const orc_sim *s, int64_t *bad);

/* O(N^2) enumeration of overlapping sphere pairs between distinct non-frozen clumps with
 * the staged test of force_part.cu:140-157 (independent of any cell grid).  Returns the
 * number of directed contacts; fills buf up to cap, sorted by (l, sm, sl, m). */
int64_t orc_brute_contacts(const avf_particle *pp, int64_t n, const avf_system *sys,
                           avf_contact *buf, int64_t cap);

#ifdef __cplusplus
}
#endif

#endif
