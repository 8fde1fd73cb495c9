/*
 * avflume.h -- C ABI of the MI355X (gfx950) DEM step for tetrahedral 4-sphere clumps.
 *
 * The reference (gsrdzl/AvalancheFlume3D) has no FFI: its operator interface is the
 * prototype set of cabecera.hu:115-157, i.e. `void op_hst(double3 *force, double3 *torque,
 * ..., particle *pp, system_par sys, ...)`, called in a fixed order by the step loops of
 * avalanche_hst.cu:438-467 (hopper filling) and avalanche_hst.cu:668-710 (avalanche).
 * This header replaces that prototype set with a context-based ABI:
 *
 *   - the POD records that cross the boundary keep the reference layout
 *     (avf_particle = particle, avf_system = system_par, avf_cell = cell_par,
 *     avf_funnel = Funnel), so a host driver written against cabecera.hu plugs in;
 *   - compile-time geometry of cabecera.hu:5-19 becomes the runtime avf_geometry;
 *   - the library owns all device memory; ops never exit(): every entry point returns
 *     an avf_status code and avf_last_error() holds the message;
 *   - one context steps one or more independent instances (a batch) in the same kernels.
 *
 * Plain pointers and sizes only; no torch or HIP types appear in the signatures.
 */
#ifndef AVFLUME_H
#define AVFLUME_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- records ---------- */

/* double3 / long3 of the reference (24 B each). */
typedef struct avf_double3 { double x, y, z; } avf_double3;
typedef struct avf_long3 { int64_t x, y, z; } avf_long3;

/* Clump record; same fields, order and size (408 B) as `particle`, cabecera.hu:77-98.
 * Rmatx/y/z are the ROWS of the body->lab rotation matrix (place_particles.cu:29-36). */
typedef struct avf_particle {
  double radius;
  double mass;
  double density;
  avf_double3 moment_of_inertia;
  avf_double3 rr_a, rr_b, rr_c, rr_d;
  avf_double3 rrot_a, rrot_b, rrot_c, rrot_d;
  avf_double3 rr;
  avf_double3 rrot;
  avf_double3 vv;
  avf_double3 ww;
  avf_double3 Rmatx, Rmaty, Rmatz;
} avf_particle;

/* System parameters; same fields, order and size (256 B) as `system_par`,
 * cabecera.hu:34-66. */
typedef struct avf_system {
  int64_t npart;
  double dt, hdt, rmin, rmax, v0, time;
  double eps_pp, eps_pw, kappa_pp, kappa_pw, gamma_pp, gamma_pw, mu_pp, mu_pw;
  double t_coll, tot_mass, tilt, xhopper, Lflume, gyr_max, fr_fact_pp, fr_fact_pw;
  double sint, cost, sintg, costg, xfunnel, dy2;
  avf_double3 side;
} avf_system;

/* Funnel parameters; `Funnel`, cabecera.hu:69-74. */
typedef struct avf_funnel { double sinfi, cosfi, tanfi; } avf_funnel;

/* Clump-cell system; same fields, order and size (168 B) as `cell_par`, cabecera.hu:101-112. */
typedef struct avf_cell {
  int64_t ntags;
  int64_t ncell3;
  avf_long3 ncell;
  int64_t ntot;
  avf_double3 cell_side, cs_inv, rr_max, rr_min, length;
} avf_cell;

/* Boundary modules evaluated by the force stage (bit mask). */
enum {
  AVF_MOD_PP = 1 << 0,       /* force_part.cu:11-266, 16 sphere pairings          */
  AVF_MOD_HOP = 1 << 1,      /* force_hop.cu:9-244, nw = 0, 1, 2                  */
  AVF_MOD_FUNNEL = 1 << 2,   /* force_funnel.cu:9-213                             */
  AVF_MOD_FCORNERS = 1 << 3, /* force_funnel.cu:216-404                           */
  AVF_MOD_WALLS = 1 << 4,    /* force_walls.cu:9-199                              */
  AVF_MOD_FLUME = 1 << 5,    /* force_flume.cu:10-182                             */
  AVF_MOD_DEP = 1 << 6       /* force_dep.cu:10-178 (never called by the reference) */
};
#define AVF_MODULES_FILL (AVF_MOD_PP | AVF_MOD_HOP)
#define AVF_MODULES_AVALANCHE \
  (AVF_MOD_HOP | AVF_MOD_FUNNEL | AVF_MOD_FCORNERS | AVF_MOD_WALLS | AVF_MOD_PP | AVF_MOD_FLUME)

/* Step composition.  FILL: avalanche_hst.cu:438-467 (pp before hop, no rotate_vectors).
 * AVALANCHE: avalanche_hst.cu:668-710 (hop, funnel, fcorners, walls, pp, flume). */
enum { AVF_PHASE_FILL = 0, AVF_PHASE_AVALANCHE = 1 };

/* Runtime form of the geometry #defines of cabecera.hu:5-19 plus the clump shape and the
 * deposit plane.  avf_geometry_reference() fills the reference values. */
typedef struct avf_geometry {
  double hop_height;  /* HOP_HEIGHT 25.0  */
  double hop_width;   /* HOP_WIDTH  15.0  */
  double yfunnel;     /* YFUNNEL    30.0  */
  double lfunnel;     /* LFUNNEL    70.0  */
  double yhop_l;      /* YHOP_L      7.5  */
  double yhop_r;      /* YHOP_R     22.5  */
  double grav;        /* GRAV      981.0  */
  double drag;        /* DRAG    1.895e-3 */
  /* Sphere centres sit at rr + (radius/sphere_div)*R*(+-1,+-1,+-1) (init_Verlet.cu:130-142).
   * Reference: ROOT2 = 1.414213562373095 (offset r*sqrt(1.5)).  Offset d: sqrt(3)*r/d. */
  double sphere_div;
  /* Clump-cell side in units of rmax: GYR + 0.05f (avalanche_hst.cu:115). */
  double cell_dist;
  /* Deposit floor z = tan_d*(x - dep_x0 - dep_x1) + dep_z0, normal (sin_d, 0, -cos_d)
   * (force_dep.cu:73-89 with tan_d = sin_d/cos_d, dep_x0 = Lflume, dep_x1 = xhopper). */
  double dep_sin, dep_cos, dep_x0, dep_x1, dep_z0;
  /* Clump-cell margin interval of cell_dim (cell_dim.cu:16, 2000 steps) and the step
   * interval of the clump-cell regrid (avalanche_hst.cu:470,713). */
  double regrid_ddt_steps;
  int32_t regrid_interval;
  int32_t inter_flag;  /* INTER_FLAG 1 (force_walls.cu:105-109)                     */
  /* 0: reference mass/inertia (calculate_particle_parameters.cu:19-33);
   * 1: mass 4 m_s, inertia 4 m_s (2/5 r^2 + 2/3 d^2) for offset d = sqrt(3) r/sphere_div. */
  int32_t clump_mode;
  int32_t modules_fill;      /* AVF_MOD_* evaluated in the fill phase                  */
  int32_t modules_avalanche; /* AVF_MOD_* evaluated in the avalanche phase             */
  int32_t reserved;
} avf_geometry;

/* One instance of a batch. */
typedef struct avf_instance_desc {
  avf_system sys;     /* derived fields filled (avf_derive_system)                      */
  avf_geometry geo;
  avf_funnel fun;
  int64_t ntags;      /* cell_par.ntags: capacity of a reference clump cell (cells.cu:96)  */
  int32_t phase;      /* AVF_PHASE_*                                                    */
  int32_t reserved;
} avf_instance_desc;

/* Library options. */
typedef struct avf_options {
  int32_t reorder_interval; /* spatial reorder of clump storage every N steps (0: never) */
  int32_t use_graph;        /* capture each reorder chunk as a hipGraph and replay it    */
  double sphere_cell_factor;/* sphere-cell side = factor * 2 * max sphere radius (>= 1)  */
  int32_t block_clumps;     /* lanes per block of the force kernel: 64 or 128          */
  int32_t max_neighbours;   /* informational: touching-clump list capacity (32)        */
  int32_t device;           /* HIP device ordinal                                      */
  int32_t lean_state;       /* candidate-list chunks: 0 or 1 = inside a chunk a clump's */
                            /* state (centre, velocities, rotation, acceleration) lives */
                            /* only in its stage record, the angular acceleration in   */
                            /* the state arrays (one read and one write of the state  */
                            /* per step); -1 = the whole state also goes to the state  */
                            /* arrays every step; same results                         */
  void *stream;             /* hipStream_t to run on; NULL = library-owned stream      */
  double verlet_skin;       /* > 0: candidate lists of sphere pairs within sep + skin, */
                            /* rebuilt only when some sphere has moved skin / 2 since  */
                            /* the build (results unchanged); 0: rebuild every step    */
  int32_t closed_forms;     /* clump mass and inertia in the step: 0 or 1 = recomputed  */
                            /* from radius and density when every clump of an instance  */
                            /* holds the closed forms bit for bit (fewer loads), -1 =   */
                            /* always loaded; same results                              */
  int32_t step_binning;     /* per-step binning (verlet_skin 0): 0 = one key per clump   */
                            /* in cells of the largest clump's bounding diameter and a  */
                            /* list of candidate sphere pairings per clump, re-tested  */
                            /* in FP64 by the force stage; 1 = four keys per clump in  */
                            /* sphere cells and per-sphere lists; same results          */
} avf_options;

/* Status codes.  Nothing in the library calls exit(). */
typedef enum avf_status {
  AVF_OK = 0,
  AVF_ERR_ARG = 1,          /* bad argument / size                                      */
  AVF_ERR_HIP = 2,          /* HIP runtime error                                        */
  AVF_ERR_NOMEM = 3,        /* allocation failed or capacity exceeded                   */
  AVF_ERR_CELL_NEG = 4,     /* a clump below the clump-cell grid (cells.cu:79-88)       */
  AVF_ERR_CELL_FULL = 5,    /* a clump cell holds ntags-1 clumps (cells.cu:96-102)      */
  AVF_ERR_NEIGH_OVERFLOW = 6,/* a clump touched more clumps than max_neighbours          */
  AVF_ERR_NOFIT = 7,        /* lattice does not fit the box (place_particles.cu:58-62)  */
  AVF_ERR_INPUT = 8         /* malformed DataInput.dat / checkpoint; unsupported state  */
} avf_status;

/* Diagnostics accumulated by the device since the last avf_get_diag(). */
typedef struct avf_diag {
  int64_t cell_below;       /* clumps whose reference clump-cell index was < 0        */
  int64_t cell_beyond;      /* clumps with a clump-cell index > ncell (reference aliases) */
  int64_t neigh_overflow;   /* clumps whose neighbour list overflowed                  */
  int64_t max_neighbours;   /* largest neighbour count seen                             */
  int64_t max_cell_spheres; /* largest sphere-cell occupancy at the last binning        */
  int64_t steps;            /* steps taken by the context                               */
  int64_t list_builds;      /* candidate-list builds                                    */
  int64_t steps_undone;     /* candidate-list steps undone and redone after a rebuild    */
                            /* (a sphere moved more than 0.49 skin within their chunk)   */
  int64_t closed_instances; /* instances whose list steps recompute clump mass and     */
                            /* inertia from radius and density (avf_options.closed_forms */
                            /* and the upload check); the others load them              */
  int64_t sweep_global_bricks; /* sphere-cell sweep bricks (initial force pass,         */
                            /* step_binning = 1, exports) whose halo did not fit LDS     */
                            /* (swept from global memory)                                */
  int64_t sweep_fp64_pairs; /* sphere pairs within the float test's margin of sep,       */
                            /* decided by the FP64 test                                  */
} avf_diag;

/* Directed sphere contact (l <- m) of force_part.cu:157-209. */
typedef struct avf_contact {
  int64_t l, m;             /* reference clump indices                                  */
  int32_t sl, sm;           /* sphere indices 0..3 (a..d) of l and m                     */
  int32_t repulsive;        /* normal > 0 (force_part.cu:209)                           */
  int32_t reserved;
} avf_contact;

typedef struct avf_ctx avf_ctx;

/* ------------------------------------------------------ host-side setup (CPU) ------ */

/* Reference geometry constants of cabecera.hu:5-19. */
void avf_geometry_reference(avf_geometry *geo);
/* Library defaults. */
void avf_options_default(avf_options *opt);

/* Lagged-Fibonacci (250,103) minus 64-bit LCG generator of rannew64.cu:55-126, with
 * its static state made explicit so instances can be seeded independently. */
typedef struct avf_rng {
  uint64_t useed[250];
  uint64_t uuy;
  int32_t np, nq, npp, not_init;
} avf_rng;
void avf_rng_init(avf_rng *st);
double avf_rannew64(avf_rng *st, int64_t *idum);

/* Parses the 16 positional lines of DataInput.dat (avalanche_hst.cu:34-81). */
int avf_parse_datainput(const char *text, avf_system *sys, int64_t *ntags, int32_t *print_flag);
/* Derived constants of avalanche_hst.cu:84-103 and :187-195 (needs npart, side.y, side.z,
 * rmin, rmax, v0, dt, eps_*, mu_*, t_coll, Lflume, tilt in degrees). */
void avf_derive_system(avf_system *sys, avf_funnel *fun, const avf_geometry *geo);
/* Input validation of error_check.cu:3-65; returns the error number (0 = ok). */
int avf_error_check(const avf_system *sys, char *msg, int msglen);
/* generate_particles.cu:7-32 (density, bidisperse radius). */
void avf_generate_particles(avf_particle *pp, const avf_system *sys, double density);
/* calculate_particle_parameters.cu:10-41 (clump_mode 0) or the general offset form. */
void avf_calculate_particle_parameters(avf_particle *pp, const avf_system *sys,
                                       const avf_geometry *geo);
/* place_particles.cu:8-90: cubic lattice in the hopper, identity orientation. */
int avf_place_particles(avf_particle *pp, const avf_system *sys, const avf_geometry *geo);
/* set_initial_velocity.cu:7-42 (seed 26752658 unless idum0 != 0). */
void avf_set_initial_velocity(avf_particle *pp, const avf_system *sys, avf_rng *st,
                              int64_t idum0);
/* Jittered lattice in the box [lo, hi] with seeded uniform-random orientations and zero
 * velocity; spacing a, jitter +-jit per axis.  Returns AVF_ERR_NOFIT if n clumps do not fit. */
int avf_place_lattice(avf_particle *pp, int64_t n, const avf_geometry *geo,
                      avf_double3 lo, avf_double3 hi, double a, double jit, int64_t seed);
/* Sphere centres rr_a..d from rr, R, radius (avalanche_hst.cu:265-293). */
void avf_sphere_centres(avf_particle *pp, int64_t n, const avf_geometry *geo);
/* cell_dim.cu:7-66 plus the regrid of avalanche_hst.cu:311-333. */
void avf_cell_setup(avf_cell *cell, const avf_particle *pp, const avf_system *sys,
                    const avf_geometry *geo);

/* ------------------------------------------------------------ device context ------- */

/* Creates a context stepping `ninst` instances of sizes inst[i].sys.npart. */
int avf_create(avf_ctx **out, const avf_instance_desc *inst, int32_t ninst,
               const avf_options *opt);
int avf_destroy(avf_ctx *ctx);
/* Message of the last failing call on ctx (or of the last failing avf_create if ctx is NULL). */
const char *avf_last_error(const avf_ctx *ctx);

/* Clump state of instance `inst` in the reference AoS layout; aa_i / anga_i are the
 * accelerations carried between steps (avalanche_hst.cu:589-595).  Either may be NULL. */
int avf_upload(avf_ctx *ctx, int32_t inst, const avf_particle *pp, const avf_double3 *aa_i,
               const avf_double3 *anga_i);
int avf_download(avf_ctx *ctx, int32_t inst, avf_particle *pp, avf_double3 *aa_i,
                 avf_double3 *anga_i);
/* Per-instance scalars that the host driver changes between phases
 * (avalanche_hst.cu:606-613: side.x = xfinal, time = 0, cc = 0). */
int avf_set_phase(avf_ctx *ctx, int32_t inst, int32_t phase, double time, double side_x,
                  int64_t cc);
int avf_get_time(avf_ctx *ctx, int32_t inst, double *time, int64_t *cc);
/* Reference clump-cell grid (cell_dim + regrid, avalanche_hst.cu:311-333) on device. */
int avf_regrid(avf_ctx *ctx);
int avf_get_cell(avf_ctx *ctx, int32_t inst, avf_cell *cell);
/* Initial force pass of avalanche_hst.cu:365-384: forces on the current state, then
 * accelerations into aa_i / anga_i. */
int avf_initial_forces(avf_ctx *ctx);
/* nsteps full steps of the instances' phase composition, including the clump-cell
 * regrid every regrid_interval steps (avalanche_hst.cu:470-505, 713-748). */
int avf_step(avf_ctx *ctx, int64_t nsteps);
/* Force and torque of the last force stage (before get_accel). */
int avf_get_forces(avf_ctx *ctx, int32_t inst, avf_double3 *force, avf_double3 *torque);
/* Reference clump-cell index per clump, cells.cu:69-91 (multiply form; -1: frozen). */
int avf_get_clump_cells(avf_ctx *ctx, int32_t inst, int64_t *cell_index);
/* Reproduces cell_locate's failure checks (cells.cu:79-102) on the current state:
 * AVF_ERR_CELL_NEG / AVF_ERR_CELL_FULL with the offending clump in *bad. */
int avf_check_cells(avf_ctx *ctx, int32_t inst, int64_t *bad);
/* Directed contacts of the current state (the set force_part visits), sorted by
 * (l, sm, sl, m).  *count receives the total even when it exceeds cap. */
int avf_get_contacts(avf_ctx *ctx, int32_t inst, avf_contact *out, int64_t cap,
                     int64_t *count);
/* The reference's clump-cell table of the current state, as clear_cell / clear_ocup /
 * cell_locate leave it (cells.cu:11-111; replaces the long *cell_vec, int *nocup arguments
 * of cabecera.hu:115-157): cell_vec[ncell3 * ntags] (-1 = empty slot, else the clump index,
 * clumps in ascending order) and nocup[ncell3]; frozen clumps are skipped.  ncell3 must be
 * the product of avf_get_cell's ncell.  Where cell_locate would exit (or, for an index past
 * the table, write out of bounds): AVF_ERR_CELL_NEG / AVF_ERR_CELL_FULL, clump in *bad. */
int avf_get_cell_table(avf_ctx *ctx, int32_t inst, int64_t *cell_vec, int32_t *nocup,
                       int64_t ncell3, int64_t *bad);
/* Sphere-cell grid of an instance: cell (i, j, k) holds the sphere centres q with
 * i = clamp(floor((q.x - lo[0]) * inv_h), 0, n[0] - 1) etc.; key = i + n[0] (j + n[1] k). */
typedef struct avf_sphere_grid {
  double lo[3];
  double h, inv_h;
  int32_t n[3];
  int32_t reserved;
} avf_sphere_grid;
/* Debug export of the binning (no reference counterpart: the reference bins clumps, not
 * spheres): the sphere-cell key of sphere s of clump m at keys[4 m + s] (0xffffffff: not
 * binned, the clump was frozen), the counting sort's order as sphere ids 4 m + s sorted by
 * key (within a cell by id: the device places a cell's spheres in arrival order, on which
 * no result depends), *nbinned entries, and the grid.  Needs a binned state (after
 * avf_initial_forces or a per-step-binning step): AVF_ERR_ARG otherwise. */
int avf_get_sphere_keys(avf_ctx *ctx, int32_t inst, uint32_t *keys, int64_t *order,
                        int64_t *nbinned, avf_sphere_grid *grid);
int avf_get_diag(avf_ctx *ctx, avf_diag *diag);

/* Observables of one instance, reduced on the device in reference clump order (so they do
 * not depend on the storage order).  vmean / wmean: mean |v| and |w| over all clumps
 * (avalanche_hst.cu:551-558, 573-574); xcm / zcm: centre of mass
 * (avalanche_hst.cu:819-829).  Runout and the deposit profile are defined in the
 * slope frame (x' along the slope, rotate_vectors.cu:28-30) over live clumps. */
typedef struct avf_observables {
  double vmean, wmean;
  double xcm, zcm;
  double runout_mean;   /* mass-weighted mean x' of the live clumps                        */
  double runout_front;  /* largest x' of a live clump                                      */
  double mass_live;     /* mass of the live clumps                                         */
  int64_t live;         /* clumps not past side.x                                          */
} avf_observables;

/* Deposit thickness profile along the slope: bin b (centre c_b = x0 + (b + 0.5)(x1-x0)/nbins)
 * holds sum_i V_i G(x'_i - c_b) / width with V_i the clump's solid volume (mass / density)
 * and G a normalised Gaussian of standard deviation sigma (0: the bin width), i.e. the
 * height of the deposited material smoothed over sigma. */
typedef struct avf_profile_spec {
  double x0, x1, width, sigma;
  int32_t nbins, reserved;
} avf_profile_spec;

/* Fills *obs and, if spec and profile are given, profile[0..nbins). */
int avf_observe(avf_ctx *ctx, int32_t inst, avf_observables *obs, const avf_profile_spec *spec,
                double *profile);
/* avf_observe with no device allocation.  No counterpart in the reference (its statistics
 * are host loops); for callers that cannot allocate device memory: monitoring loops of a
 * serving process, code run next to graph capture, harnesses that forbid hipMalloc.  Same
 * kernels, grid and results as avf_observe; the block partial sums live in a fixed device
 * array of 1,048,576 doubles (8 MiB) on each GPU, held until the process ends and shared by
 * every call on that GPU, so calls that overlap on one GPU race.  AVF_ERR_NOMEM when
 * (ceil(npart / 256) + 1) * (9 + nbins) exceeds it (e.g. above 3.6M clumps with 64 bins). */
int avf_observe_noalloc(avf_ctx *ctx, int32_t inst, avf_observables *obs,
                        const avf_profile_spec *spec, double *profile);

/* Kernels of one step, in launch order (indices of avf_step_profiled's kernel_ms).  Slots
 * of kernels a step does not launch read as empty (near 0): the binning kernels on
 * candidate-list steps, the predict after the first step of a chunk (it runs inside the
 * previous step's force kernel), the regrid reduction (inside the force kernel), the
 * finalize before the last step of a chunk, and the next step's cell counts (per-step
 * binning by sphere only, after every step of a chunk but its last; binned by clump, the
 * force kernel counts the next step's keys itself).  On per-step binning SCAN_A (the count
 * sums) runs on a chunk's first step only (later steps add the sums up while counting) and
 * SCAN_C is always empty. */
enum {
  AVF_K_PREDICT, AVF_K_SCAN_A, AVF_K_SCAN_B, AVF_K_SCAN_C, AVF_K_SCATTER, AVF_K_DETECT,
  AVF_K_FORCE, AVF_K_CELLREDUCE, AVF_K_FINALIZE, AVF_K_COUNT, AVF_NKERNELS
};
/* Like avf_step, but launches eagerly with HIP events between consecutive kernels on the
 * context's stream: kernel_ms[k] receives the mean duration of kernel k per step and
 * *total_ms the elapsed time from the first to the last kernel of the run. */
int avf_step_profiled(avf_ctx *ctx, int64_t nsteps, double *kernel_ms, double *total_ms);
/* Like avf_step_profiled with events around one kernel only (AVF_K_*), on every
 * `every`-th step (1: all).  For AVF_K_FORCE on candidate-list steps with every > 1, one
 * pair of events brackets a run of `every` consecutive k_force launches (nothing else runs
 * between them), one run in every 2 * `every` steps, so the events' own cost is shared by
 * the run.  *kernel_ms receives the mean duration per timed launch, *total_ms the elapsed
 * time of the run. */
int avf_step_timed(avf_ctx *ctx, int64_t nsteps, int32_t kernel, int32_t every,
                   double *kernel_ms, double *total_ms);
/* Waits for all work queued on the context's stream. */
int avf_sync(avf_ctx *ctx);

/* ---------------------------------------------------------------- differentiable window
 * No counterpart in the reference: the adjoint of its step (contact law of
 * force_part.cu:119-242 and the boundary modules, get_accel.cu, init_/finish_Verlet.cu)
 * over a window of steps with the contact list frozen at the window start.
 *
 * Contact-law parameters per instance, particle-particle (pp) and particle-wall (pw):
 * friction mu, restitution eps and normal stiffness kappa = 2 k_n / m (the reference's
 * kappa_pp / kappa_pw).  The damping rate follows the reference's contact-time relations
 * (avalanche_hst.cu:187-195) with kappa held fixed: t_c = sqrt((pi^2 + ln^2 eps) / kappa),
 * gamma = -2 ln(eps) / t_c.  theta[inst * AVF_NTHETA + k]. */
enum {
  AVF_THETA_MU_PP, AVF_THETA_EPS_PP, AVF_THETA_KAPPA_PP,
  AVF_THETA_MU_PW, AVF_THETA_EPS_PW, AVF_THETA_KAPPA_PW, AVF_NTHETA
};

typedef struct avf_window_spec {
  int64_t steps;             /* window length W (steps)                                    */
  int64_t checkpoint;        /* K: state checkpoint every K steps; the backward pass       */
                             /* recomputes each K-step segment from its checkpoint         */
  double skin;               /* sphere pairs closer than r_l + r_m + skin at the window    */
                             /* start are listed; no other pair interacts in the window   */
  double w_runout;           /* loss = w_runout * runout_mean                              */
  double w_profile;          /*      + w_profile * sum_b (h_b - target_b)^2, per instance */
  avf_profile_spec profile;  /* bins of h (nbins 0: no profile term)                      */
} avf_window_spec;

typedef struct avf_window_diag {
  double max_disp;      /* largest sphere displacement from the window start                */
  int64_t max_list;     /* largest listed pair count of a clump                             */
  int64_t max_in;       /* largest count of entries naming one clump                        */
  int64_t missed;       /* touching sphere pairs absent from the list, summed over the      */
                        /* checkpoint states and the final state                            */
  int32_t list_valid;   /* missed == 0                                                      */
  int32_t reserved;
  double forward_ms;    /* list build + forward window + loss (HIP events)                  */
  double tangent_ms;    /* dual-number forward window (0 when not requested)                */
  double backward_ms;   /* segment recompute + reverse steps + gradient reductions          */
} avf_window_diag;

/* Runs the window from the context's current state (rr, R, v, w and the stored
 * accelerations a_i, alpha_i of avf_upload / the last step), which it leaves unchanged.
 * The instances must be in the avalanche phase at time > 0, with isotropic clump inertia,
 * and without the funnel / funnel-corner modules.
 *   theta      [ninst * AVF_NTHETA] or NULL (the instances' own parameters);
 *   dv         [nclumps * 3] start-velocity perturbation (m/s) or NULL; clump arrays are
 *              indexed (first clump of the instance + reference index) * 3 + axis;
 *   target     [ninst * nbins] target profile or NULL (zero);
 *   loss       [ninst] loss per instance; obs [ninst] / profile [ninst * nbins] optional
 *              observables of the final state;
 *   grad_theta [ninst * AVF_NTHETA], grad_dv [nclumps * 3]: dloss/dtheta, dloss/ddv
 *              (the reverse pass runs if either is non-NULL);
 *   tan_theta, tan_dv, dloss: if dloss is non-NULL, a dual-number forward along the
 *              direction (tan_theta, tan_dv) (NULL = 0) gives dloss[inst] = directional
 *              derivative of loss[inst];
 *   final_state [nclumps * 24] optional: final rr, vv, ww, R (row-major), a, alpha.
 * Returns AVF_ERR_INPUT for an unsupported state (message in avf_last_error) and
 * AVF_ERR_NEIGH_OVERFLOW when a clump has more listed pairs than the list holds. */
int avf_window_run(avf_ctx *ctx, const avf_window_spec *spec, const double *theta,
                   const double *dv, const double *target, double *loss, avf_observables *obs,
                   double *profile, double *grad_theta, double *grad_dv, const double *tan_theta,
                   const double *tan_dv, double *dloss, double *final_state,
                   avf_window_diag *diag);

#ifdef __cplusplus
}
#endif

#endif /* AVFLUME_H */
