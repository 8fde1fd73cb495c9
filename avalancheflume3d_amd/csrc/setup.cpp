// Host-side setup of the DEM step: the reference's DataInput.dat-driven configuration,
// particle generation and placement, restated in C++ behind the C ABI of avflume.h.
// These run on the CPU once per run; the step itself runs in the gfx950 kernels.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "avflume.h"

namespace {

// Literals of cabecera.hu:5-19, as written (ROOT2 is 1 ulp below sqrt(2)).
constexpr double kPi = 3.1415926535897932;
constexpr double kGyr = 4.4494897427831781;
constexpr double kRoot2 = 1.414213562373095;
constexpr double kGrav = 981.0;
constexpr double kDrag = 1.895e-3;

constexpr int kNP = 250;  // rannew64.cu:52-53 (Coddington taps 250, 103)
constexpr int kNQ = 103;

void set_msg(char *msg, int len, const char *text) {
  if (msg && len > 0) {
    std::snprintf(msg, (size_t)len, "%s", text);
  }
}

}  // namespace

extern "C" {

void avf_geometry_reference(avf_geometry *g) {
  std::memset(g, 0, sizeof(*g));
  g->hop_height = 25.0;
  g->hop_width = 15.0;
  g->yfunnel = 30.0;
  g->lfunnel = 70.0;
  g->yhop_l = 7.5;
  g->yhop_r = 22.5;
  g->grav = kGrav;
  g->drag = kDrag;
  g->sphere_div = kRoot2;
  g->cell_dist = kGyr + 0.05f;  // float literal, avalanche_hst.cu:115
  g->dep_sin = 0.0;
  g->dep_cos = 1.0;
  g->regrid_ddt_steps = 2000.0;
  g->regrid_interval = 2000;
  g->inter_flag = 1;
  g->clump_mode = 0;
  g->modules_fill = AVF_MODULES_FILL;
  g->modules_avalanche = AVF_MODULES_AVALANCHE;
}

void avf_options_default(avf_options *o) {
  std::memset(o, 0, sizeof(*o));
  o->reorder_interval = 50;
  o->use_graph = 1;
  o->sphere_cell_factor = 1.0;
  o->block_clumps = 64;
  o->max_neighbours = 32;
  o->device = 0;
  o->stream = nullptr;
}

// ------------------------------------------------------------------------ rannew64 ----

void avf_rng_init(avf_rng *st) {
  std::memset(st, 0, sizeof(*st));
  st->np = kNP - 1;  // static initialisers of rannew64.cu:63-64, never reset afterwards
  st->nq = kNQ - 1;
  st->not_init = 1;
}

// rannew64.cu:55-126: x(n) = x(n-250) - x(n-103), y(n) = a y(n-1) + c, z = x - y (mod 2^64).
// (Re)initialised on the first call and whenever *idum < 0.
double avf_rannew64(avf_rng *st, int64_t *pidum) {
  const uint64_t ua = 2862933555777941757ull, uc = 3037000493ull;
  const double xmm = 18446744073709551615.;
  int64_t idum = *pidum;
  if (st->not_init || idum < 0) {
    const uint64_t uaa = 6364136223846793005ull, ucc = 1442695040888963407ull;
    if (idum < 0) {
      idum = -idum;
      *pidum = idum;
    }
    uint64_t udum = (uint64_t)idum;
    for (int i = 0; i < 1000; i++) udum = udum * uaa + ucc;
    for (int i = 0; i < kNP; i++) {
      udum = udum * uaa + ucc;
      st->useed[i] = udum;
    }
    st->npp = kNP - 1;
    st->uuy = udum;
    st->not_init = 0;
  }
  uint64_t uux = st->useed[st->np] - st->useed[st->nq];
  st->useed[st->np] = uux;
  st->np--;
  st->nq--;
  if (st->np < 0) st->np = st->npp;
  if (st->nq < 0) st->nq = st->npp;
  st->uuy = ua * st->uuy + uc;
  uint64_t uu = uux - st->uuy;
  const double omm = 1. / xmm;
  return omm * ((double)uu);
}

// ------------------------------------------------------------------- DataInput.dat ----

// avalanche_hst.cu:34-81: 16 lines, the first token of each is the value.
int avf_parse_datainput(const char *text, avf_system *sys, int64_t *ntags, int32_t *print_flag) {
  if (!text || !sys) return AVF_ERR_ARG;
  std::memset(sys, 0, sizeof(*sys));
  const char *p = text;
  char row[128];
  double *dst[] = {&sys->side.y, &sys->side.z, &sys->rmin, &sys->rmax, &sys->v0, &sys->dt,
                   &sys->eps_pw, &sys->mu_pw, &sys->eps_pp, &sys->mu_pp, &sys->t_coll,
                   &sys->Lflume, &sys->tilt};
  for (int line = 0; line < 16; line++) {
    if (!*p) return AVF_ERR_INPUT;
    const char *e = std::strchr(p, '\n');
    size_t len = e ? (size_t)(e - p) : std::strlen(p);
    if (len >= sizeof(row)) len = sizeof(row) - 1;
    std::memcpy(row, p, len);
    row[len] = 0;
    p = e ? e + 1 : p + len;
    int ok;
    if (line == 0) {
      long v;
      ok = std::sscanf(row, "%ld", &v) == 1;
      sys->npart = v;
    } else if (line <= 13) {
      ok = std::sscanf(row, "%lf", dst[line - 1]) == 1;
    } else if (line == 14) {
      long v;
      ok = std::sscanf(row, "%ld", &v) == 1;
      if (ntags) *ntags = v;
    } else {
      int v;
      ok = std::sscanf(row, "%d", &v) == 1;
      if (print_flag) *print_flag = v;
    }
    if (!ok) return AVF_ERR_INPUT;
  }
  return AVF_OK;
}

// avalanche_hst.cu:84-103 and :187-195.  tot_mass (avalanche_hst.cu:181-183) needs the
// particles and is left to the caller.
void avf_derive_system(avf_system *sys, avf_funnel *fun, const avf_geometry *geo) {
  sys->side.x = geo->hop_width;
  sys->time = -1.0;
  sys->hdt = 0.5 * sys->dt;
  sys->tilt *= kPi / 180.0;
  sys->gyr_max = kGyr * sys->rmax;
  sys->sint = std::sin(sys->tilt);
  sys->cost = std::cos(sys->tilt);
  sys->sintg = geo->grav * sys->sint;
  sys->costg = geo->grav * sys->cost;
  sys->xfunnel = geo->lfunnel * sys->cost;
  sys->dy2 = 0.5 * (geo->yfunnel - sys->side.y);
  if (fun) {
    double hipof = std::sqrt(sys->dy2 * sys->dy2 + geo->lfunnel * geo->lfunnel);
    fun->sinfi = sys->dy2 / hipof;
    fun->cosfi = geo->lfunnel / hipof;
    fun->tanfi = sys->dy2 / geo->lfunnel;
  }
  double aux_0 = kPi / sys->t_coll;
  double aux_1 = std::log(sys->eps_pp) / sys->t_coll;
  sys->gamma_pp = -2.0 * std::log(sys->eps_pp) / sys->t_coll;
  sys->kappa_pp = (aux_0 * aux_0 + aux_1 * aux_1);
  aux_1 = std::log(sys->eps_pw) / sys->t_coll;
  sys->gamma_pw = -2.0 * std::log(sys->eps_pw) / sys->t_coll;
  sys->kappa_pw = (aux_0 * aux_0 + aux_1 * aux_1);
  sys->fr_fact_pp = sys->mu_pp / 1.0e12;
  sys->fr_fact_pw = sys->mu_pw / 1.0e12;
}

// error_check.cu:3-65; the last failing check sets the number, as in the reference.
int avf_error_check(const avf_system *s, char *msg, int len) {
  int errn = 0;
  char buf[160];
  if (s->npart <= 0) {
    std::snprintf(buf, sizeof buf, "ERROR. No particles in box! n = %ld\n", (long)s->npart);
    set_msg(msg, len, buf);
    errn = 1;
  }
  if (s->side.x <= 0.0 || s->side.y <= 0.0 || s->side.z <= 0.0) {
    std::snprintf(buf, sizeof buf, "ERROR. Wrong box dimensions. x = %lf, y = %lf, z = %lf\n",
                  s->side.x, s->side.y, s->side.z);
    set_msg(msg, len, buf);
    errn = 2;
  }
  if (s->rmin <= 0.0 || s->rmax <= 0.0) {
    std::snprintf(buf, sizeof buf, "ERROR. Wrong sphere radius. rmax = %lf, rmin = %lf\n",
                  s->rmax, s->rmin);
    set_msg(msg, len, buf);
    errn = 3;
  }
  if (s->rmax < s->rmin) {
    std::snprintf(buf, sizeof buf, "ERROR rmax must be larger than rmin. rmax = %lf, rmin = %lf\n",
                  s->rmax, s->rmin);
    set_msg(msg, len, buf);
    errn = 4;
  }
  if (s->eps_pw <= 0.0 || s->eps_pw > 1.0) {
    set_msg(msg, len, "ERROR. eps_pw must be 0 < eps_pw <= 1\n");
    errn = 5;
  }
  if (s->eps_pp <= 0.0 || s->eps_pp > 1.0) {
    set_msg(msg, len, "ERROR. eps_pp must be 0 < eps_pp <= 1\n");
    errn = 6;
  }
  if (s->mu_pw < 0.0 || s->mu_pw > 1.0) {
    set_msg(msg, len, "ERROR. mu_pw must be 0 < mu_pw < 1\n");
    errn = 7;
  }
  if (s->mu_pp < 0.0 || s->mu_pp > 1.5) {
    set_msg(msg, len, "ERROR. mu_pp must be 0 < mu_pp < 1\n");
    errn = 8;
  }
  if (std::tan(s->tilt) < 1.25 * s->mu_pw) {
    set_msg(msg, len,
            "ERROR. Particles won't go down the slope.\n Please, increase tilt or reduce mu_pw.\n");
    errn = 9;
  }
  return errn;
}

// ------------------------------------------------------------ particle generation ----

// generate_particles.cu:7-32: every 20th clump has radius rmax, the rest rmin.
void avf_generate_particles(avf_particle *pp, const avf_system *sys, double density) {
  int nn = (int)sys->npart;  // `int nn` in the reference
  for (int i = 0; i < nn; i++) pp[i].density = density;
  for (int i = 0; i < nn; i++) pp[i].radius = (i % 20) ? sys->rmin : sys->rmax;
}

// calculate_particle_parameters.cu:10-41 (clump_mode 0: mass 4*(4 pi/3) r^3 rho,
// inertia 112 pi/15 rho r^5, i.e. offset r*sqrt(1.5)).  clump_mode 1: the same mass and
// I = 4 m_s (2/5 r^2 + 2/3 d^2) for the offset d = sqrt(3) r / sphere_div.
void avf_calculate_particle_parameters(avf_particle *pp, const avf_system *sys,
                                       const avf_geometry *geo) {
  int nn = (int)sys->npart;
  double aux1 = 4.0 * kPi / 3.0, aux2 = 112.0 * kPi / 15.0;
  for (int i = 0; i < nn; i++) {
    double rad = pp[i].radius, rho = pp[i].density;
    double rad3 = rad * rad * rad, rad5 = rad3 * rad * rad, I;
    double mass = 4.0 * aux1 * rad3 * rho;
    if (geo->clump_mode == 0) {
      I = aux2 * rho * rad5;
    } else {
      double ms = aux1 * rad3 * rho, d = std::sqrt(3.0) * rad / geo->sphere_div;
      I = 4.0 * ms * (0.4 * rad * rad + (2.0 / 3.0) * d * d);
    }
    pp[i].mass = mass;
    pp[i].moment_of_inertia.x = pp[i].moment_of_inertia.y = pp[i].moment_of_inertia.z = I;
  }
}

// place_particles.cu:8-90: cubic lattice (spacing (2 + ROOT2 + 0.05 rmax) rmax) filling the
// hopper x fastest, then y, then z above HOP_HEIGHT; identity orientation.
int avf_place_particles(avf_particle *pp, const avf_system *sys, const avf_geometry *geo) {
  int nn = (int)sys->npart;
  double rmax = sys->rmax;
  double dist = 2.0 + kRoot2 + 0.05 * rmax, hdist = dist / 2.0;
  int ii = 0, jj = 0, kk = 0;
  for (int a = 0; a < nn; a++) {
    avf_double3 rr;
    rr.x = hdist * rmax + (double)ii * dist * rmax;
    rr.y = hdist * rmax + (double)jj * dist * rmax + 0.5 * rmax + geo->yhop_l;
    rr.z = hdist * rmax + (double)kk * dist * rmax + 0.5 * rmax;
    ii++;
    if (rr.x >= sys->side.x - hdist * rmax) {
      rr.x = hdist * rmax;
      rr.y += dist * rmax;
      jj++;
      ii = 1;
      if (rr.y >= geo->yhop_r - hdist * rmax) {
        rr.y = hdist * rmax + 0.5 * rmax + geo->yhop_l;
        rr.z += dist * rmax;
        kk++;
        jj = 0;
        if (rr.z >= sys->side.z - hdist * rmax) return AVF_ERR_NOFIT;
      }
    }
    rr.z += geo->hop_height;
    double pos = pp[a].radius / geo->sphere_div;
    pp[a].rr = rr;
    pp[a].rr_a = {rr.x + pos, rr.y + pos, rr.z + pos};
    pp[a].rr_b = {rr.x + pos, rr.y - pos, rr.z - pos};
    pp[a].rr_c = {rr.x - pos, rr.y - pos, rr.z + pos};
    pp[a].rr_d = {rr.x - pos, rr.y + pos, rr.z - pos};
    pp[a].Rmatx = {1.0, 0.0, 0.0};
    pp[a].Rmaty = {0.0, 1.0, 0.0};
    pp[a].Rmatz = {0.0, 0.0, 1.0};
  }
  return AVF_OK;
}

// set_initial_velocity.cu:7-42: |v| = v0 in a random direction of the upper hemisphere,
// w = v/(4 rmax).  Seed 26752658 unless idum0 != 0.
void avf_set_initial_velocity(avf_particle *pp, const avf_system *sys, avf_rng *st,
                              int64_t idum0) {
  int nn = (int)sys->npart;
  int64_t idum = idum0 ? idum0 : 26752658;
  double v0 = sys->v0, rmax = sys->rmax, inv_rm = 1.0 / (4.0 * rmax);
  for (int i = 0; i < nn; i++) {
    double r0 = 2.0 * kPi * avf_rannew64(st, &idum) - kPi;
    double r1 = kPi * avf_rannew64(st, &idum);
    pp[i].vv = {v0 * std::cos(r0) * std::cos(r1), v0 * std::sin(r0) * std::cos(r1),
                v0 * std::sin(r1)};
    pp[i].ww = {inv_rm * v0 * std::cos(r0) * std::cos(r1),
                inv_rm * v0 * std::sin(r0) * std::cos(r1), inv_rm * v0 * std::sin(r1)};
  }
}

// Sphere centres from rr, R and the radius: the arithmetic of avalanche_hst.cu:265-293 and
// init_Verlet.cu:130-142 (R rows Rmatx/y/z; tetrahedron (1,1,1), (1,-1,-1), (-1,-1,1),
// (-1,1,-1)).
void avf_sphere_centres(avf_particle *pp, int64_t n, const avf_geometry *geo) {
  for (int64_t i = 0; i < n; i++) {
    avf_particle &p = pp[i];
    double pos = p.radius / geo->sphere_div;
    const avf_double3 &rr = p.rr, &Rx = p.Rmatx, &Ry = p.Rmaty, &Rz = p.Rmatz;
    p.rr_a = {rr.x + pos * (Rx.x + Rx.y + Rx.z), rr.y + pos * (Ry.x + Ry.y + Ry.z),
              rr.z + pos * (Rz.x + Rz.y + Rz.z)};
    p.rr_b = {rr.x + pos * (Rx.x - Rx.y - Rx.z), rr.y + pos * (Ry.x - Ry.y - Ry.z),
              rr.z + pos * (Rz.x - Rz.y - Rz.z)};
    p.rr_c = {rr.x + pos * (-Rx.x - Rx.y + Rx.z), rr.y + pos * (-Ry.x - Ry.y + Ry.z),
              rr.z + pos * (-Rz.x - Rz.y + Rz.z)};
    p.rr_d = {rr.x + pos * (-Rx.x + Rx.y - Rx.z), rr.y + pos * (-Ry.x + Ry.y - Ry.z),
              rr.z + pos * (-Rz.x + Rz.y - Rz.z)};
  }
}

// Jittered lattice with uniform-random orientations (unit quaternions by Shoemake's
// subgroup method from three uniforms of rannew64), zero velocity.  Radii must be set.
int avf_place_lattice(avf_particle *pp, int64_t n, const avf_geometry *geo, avf_double3 lo,
                      avf_double3 hi, double a, double jit, int64_t seed) {
  if (n < 0 || a <= 0.0) return AVF_ERR_ARG;
  // counts per axis; the 1e-9 guards box lengths that are exact multiples of a
  int64_t nx = (int64_t)((hi.x - lo.x) / a + 1e-9), ny = (int64_t)((hi.y - lo.y) / a + 1e-9);
  int64_t nz = (int64_t)((hi.z - lo.z) / a + 1e-9);
  if (nx < 1 || ny < 1 || nz < 1 || nx * ny * nz < n) return AVF_ERR_NOFIT;
  avf_rng st;
  avf_rng_init(&st);
  int64_t idum = -(seed ? (seed < 0 ? -seed : seed) : 1);
  for (int64_t i = 0; i < n; i++) {
    int64_t ix = i % nx, iy = (i / nx) % ny, iz = i / (nx * ny);
    avf_particle &p = pp[i];
    double j0 = avf_rannew64(&st, &idum), j1 = avf_rannew64(&st, &idum);
    double j2 = avf_rannew64(&st, &idum);
    p.rr = {lo.x + (ix + 0.5) * a + jit * (2.0 * j0 - 1.0),
            lo.y + (iy + 0.5) * a + jit * (2.0 * j1 - 1.0),
            lo.z + (iz + 0.5) * a + jit * (2.0 * j2 - 1.0)};
    double u1 = avf_rannew64(&st, &idum), u2 = avf_rannew64(&st, &idum);
    double u3 = avf_rannew64(&st, &idum);
    double s1 = std::sqrt(1.0 - u1), s2 = std::sqrt(u1);
    double qw = s1 * std::sin(2.0 * kPi * u2), qx = s1 * std::cos(2.0 * kPi * u2);
    double qy = s2 * std::sin(2.0 * kPi * u3), qz = s2 * std::cos(2.0 * kPi * u3);
    p.Rmatx = {1.0 - 2.0 * (qy * qy + qz * qz), 2.0 * (qx * qy - qz * qw),
               2.0 * (qx * qz + qy * qw)};
    p.Rmaty = {2.0 * (qx * qy + qz * qw), 1.0 - 2.0 * (qx * qx + qz * qz),
               2.0 * (qy * qz - qx * qw)};
    p.Rmatz = {2.0 * (qx * qz - qy * qw), 2.0 * (qy * qz + qx * qw),
               1.0 - 2.0 * (qx * qx + qy * qy)};
    p.vv = {0.0, 0.0, 0.0};
    p.ww = {0.0, 0.0, 0.0};
  }
  avf_sphere_centres(pp, n, geo);
  return AVF_OK;
}

// cell_dim.cu:7-66 plus the regrid arithmetic of avalanche_hst.cu:311-333.  cell->ntags
// must be set by the caller.
void avf_cell_setup(avf_cell *cell, const avf_particle *pp, const avf_system *sys,
                    const avf_geometry *geo) {
  int nn = (int)sys->npart;
  double rgyr = 0.5 * sys->gyr_max, ddt = geo->regrid_ddt_steps * sys->dt;
  avf_double3 rmax = pp[0].rr, rmin = pp[0].rr;
  double vxmax = pp[0].vv.x, vzmin = pp[0].vv.z;
  for (int i = 1; i < nn; i++) {
    const avf_double3 &rr = pp[i].rr, &vv = pp[i].vv;
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
  cell->rr_max = rmax;
  cell->rr_min = rmin;
  double dist = geo->cell_dist;
  cell->length = {rmax.x - rmin.x, rmax.y - rmin.y, rmax.z - rmin.z};
  cell->ncell.x = (int64_t)(cell->length.x / (dist * sys->rmax));
  cell->ncell.y = (int64_t)(cell->length.y / (dist * sys->rmax));
  cell->ncell.z = (int64_t)(cell->length.z / (dist * sys->rmax));
  cell->cell_side.x = cell->length.x / ((double)cell->ncell.x);
  cell->cell_side.y = cell->length.y / ((double)cell->ncell.y);
  cell->cell_side.z = cell->length.z / ((double)cell->ncell.z);
  cell->cs_inv = {1.0 / cell->cell_side.x, 1.0 / cell->cell_side.y, 1.0 / cell->cell_side.z};
  cell->ncell3 = cell->ncell.x * cell->ncell.y * cell->ncell.z;
  cell->ntot = cell->ncell3 * cell->ntags;
}

}  // extern "C"
