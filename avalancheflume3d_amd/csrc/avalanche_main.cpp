// avalanche_amd -- host driver of the gfx950 DEM step.
//
// Reads DataInput.dat on stdin and runs the two-phase experiment of avalanche_hst.cu:
// hopper filling from t = -1 s until t >= 0 (or the stale mean speed drops below v0/10),
// gate opening (side.x = xfinal, t = 0), then the avalanche until t = 2 s (or the mean
// speed drops below v0/10 after t = 0.2 s).  Every step runs on the GPU through the C ABI;
// the clump-cell regrid every 2,000 steps happens on the device, and the state comes back to
// the host only every 20,000 steps for the reference's outputs:
//   log_amd.dat, mean_velocity_amd.csv, xcm_amd.csv, snapshots_amd.xyz, velocities_amd.xyz,
//   last_config.dat (the reference's 18-column text checkpoint) and last_config.bin (lossless).
//
//   avalanche_amd [--restart] [--restart-bin FILE] [--max-steps N] [--final-bin FILE]
//                 [--device D] [--skin CM] [--reorder N] < DataInput.dat
// --skin: candidate-list skin (default a quarter of the largest sphere radius, 0.045 cm for
// the reference's DataInput.dat; 0 bins the spheres every step), --reorder: spatial reorder
// interval (default 400 steps).  Neither changes any result.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "avflume.h"

namespace {

constexpr double kPi = 3.1415926535897932;
constexpr int64_t kRegridEvery = 2000;    // avalanche_hst.cu:470, 713
constexpr int64_t kOutputEvery = 20000;   // avalanche_hst.cu:507, 750
constexpr uint64_t kBinMagic = 0x314b504346564100ull;  // "\0AVFCPK1"

struct Outputs {
  FILE *log = nullptr, *xyz = nullptr, *vmean = nullptr, *vel = nullptr, *xcm = nullptr;
};

void die(const char *what, avf_ctx *ctx, int rc) {
  std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, avf_last_error(ctx));
  std::exit(1);
}

#define CK(call)                              \
  do {                                        \
    int rc_ = (call);                         \
    if (rc_) die(#call, ctx, rc_);            \
  } while (0)

// avalanche_hst.cu:519-560 (filling) and :763-804 (avalanche): text checkpoint, .xyz
// frames, mean speeds; xcm / zcm only during the avalanche (avalanche_hst.cu:819-829).
void write_outputs(const Outputs &o, const avf_system &sys, const std::vector<avf_particle> &pp,
                   int aval_start, int64_t cc, int64_t jj, int print_flag, double xfinal,
                   double yfunnel, double *vmean_out) {
  const int64_t nn = sys.npart;
  double vmean = 0.0, wmean = 0.0, xcm = 0.0, zcm = 0.0;
  if (print_flag) {
    std::fprintf(o.xyz, "%ld\n\n", (long)(4 * nn + 2));
    std::fprintf(o.vel, "%ld\n\n", (long)(nn + 2));
  }
  FILE *conf = std::fopen("last_config.dat", "w");
  std::fprintf(conf, "%d\n%lf\n%ld\n%ld\n", aval_start, sys.time, (long)(cc + 1), (long)(jj + 1));
  for (int64_t i = 0; i < nn; i++) {
    const avf_particle &p = pp[i];
    if (print_flag) {
      const avf_double3 *s[4] = {&p.rr_a, &p.rr_b, &p.rr_c, &p.rr_d};
      for (int k = 0; k < 4; k++)
        std::fprintf(o.xyz, "%lf %lf %lf %lf %ld\n", p.radius, s[k]->x, s[k]->y, s[k]->z, (long)i);
      std::fprintf(o.vel, "%lf %lf %lf %lf %lf %ld\n", p.rr.x, p.rr.y, p.rr.z, p.vv.x, p.ww.y,
                   (long)i);
    }
    std::fprintf(conf,
                 "%ld %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf\n",
                 (long)i, p.rr.x, p.rr.y, p.rr.z, p.Rmatx.x, p.Rmatx.y, p.Rmatx.z, p.Rmaty.x,
                 p.Rmaty.y, p.Rmaty.z, p.Rmatz.x, p.Rmatz.y, p.Rmatz.z, p.vv.x, p.vv.y, p.vv.z,
                 p.ww.x, p.ww.y, p.ww.z);
    double vm2 = p.vv.x * p.vv.x;
    vm2 += p.vv.y * p.vv.y;
    vm2 += p.vv.z * p.vv.z;
    vmean += std::sqrt(vm2);
    double wm2 = p.ww.x * p.ww.x;
    wm2 += p.ww.y * p.ww.y;
    wm2 += p.ww.z * p.ww.z;
    wmean += std::sqrt(wm2);
  }
  std::fclose(conf);
  if (print_flag) {
    if (aval_start)
      std::fprintf(o.xyz, "0.0001 %lf %lf %lf %ld\n", sys.side.x, yfunnel, sys.side.z, (long)nn);
    else
      std::fprintf(o.xyz, "0.000001 %lf %lf %lf %ld\n", xfinal, yfunnel, sys.side.z, (long)nn);
    std::fprintf(o.xyz, "0.000001 0.000000 0.000000 0.000000 %ld\n", (long)nn);
    std::fprintf(o.vel, "%lf %lf %lf 0.000000 0.000000 %ld\n", xfinal, yfunnel, sys.side.z,
                 (long)nn);
    std::fprintf(o.vel, "0.000000 0.000000 0.000000 0.000000 0.000000 %ld\n", (long)nn);
  }
  vmean /= nn;
  wmean /= nn;
  std::fprintf(o.vmean, "%lf, %lf, %lf\n", sys.time - sys.dt, vmean, wmean);
  if (aval_start) {
    for (int64_t i = 0; i < nn; i++) {
      xcm += pp[i].mass * pp[i].rr.x;
      zcm += pp[i].mass * pp[i].rr.z;
    }
    xcm /= sys.tot_mass;
    zcm /= sys.tot_mass;
    std::fprintf(o.xcm, "%lf, %lf, %lf\n", sys.time - sys.dt, xcm, zcm);
  }
  std::fflush(o.vmean);
  std::fflush(o.xcm);
  if (print_flag) {
    std::fflush(o.xyz);
    std::fflush(o.vel);
  }
  *vmean_out = vmean;
}

// Lossless checkpoint: header, the particle records and the carried accelerations.
void write_bin(const char *path, const avf_system &sys, int aval_start, int64_t cc, int64_t jj,
               const std::vector<avf_particle> &pp, const std::vector<avf_double3> &aa,
               const std::vector<avf_double3> &ga) {
  FILE *f = std::fopen(path, "wb");
  if (!f) return;
  int64_t hdr[4] = {(int64_t)aval_start, cc, jj, sys.npart};
  std::fwrite(&kBinMagic, sizeof kBinMagic, 1, f);
  std::fwrite(hdr, sizeof hdr, 1, f);
  std::fwrite(&sys, sizeof sys, 1, f);
  std::fwrite(pp.data(), sizeof(avf_particle), pp.size(), f);
  std::fwrite(aa.data(), sizeof(avf_double3), aa.size(), f);
  std::fwrite(ga.data(), sizeof(avf_double3), ga.size(), f);
  std::fclose(f);
}

bool read_bin(const char *path, avf_system *sys, int *aval_start, int64_t *cc, int64_t *jj,
              std::vector<avf_particle> &pp, std::vector<avf_double3> &aa,
              std::vector<avf_double3> &ga) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  uint64_t magic = 0;
  int64_t hdr[4];
  bool ok = std::fread(&magic, sizeof magic, 1, f) == 1 && magic == kBinMagic &&
            std::fread(hdr, sizeof hdr, 1, f) == 1 && std::fread(sys, sizeof *sys, 1, f) == 1 &&
            hdr[3] == (int64_t)pp.size();
  if (ok) {
    *aval_start = (int)hdr[0];
    *cc = hdr[1];
    *jj = hdr[2];
    ok = std::fread(pp.data(), sizeof(avf_particle), pp.size(), f) == pp.size() &&
         std::fread(aa.data(), sizeof(avf_double3), aa.size(), f) == aa.size() &&
         std::fread(ga.data(), sizeof(avf_double3), ga.size(), f) == ga.size();
  }
  std::fclose(f);
  return ok;
}

// avalanche_hst.cu:218-263: the text checkpoint (6 decimals, lossy); radius, mass and
// inertia stay those regenerated from DataInput.dat.
bool read_text(const char *path, avf_system *sys, int *aval_start, int64_t *cc, int64_t *jj,
               std::vector<avf_particle> &pp) {
  FILE *f = std::fopen(path, "r");
  if (!f) return false;
  long c, j;
  bool ok = std::fscanf(f, "%d\n", aval_start) == 1 && std::fscanf(f, "%lf\n", &sys->time) == 1 &&
            std::fscanf(f, "%ld\n", &c) == 1 && std::fscanf(f, "%ld\n", &j) == 1;
  *cc = c;
  *jj = j;
  long ii;
  while (ok && std::fscanf(f, "%ld ", &ii) == 1) {
    if (ii < 0 || ii >= (long)pp.size()) {
      ok = false;
      break;
    }
    avf_particle &p = pp[ii];
    int n = std::fscanf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf\n",
                        &p.rr.x, &p.rr.y, &p.rr.z, &p.Rmatx.x, &p.Rmatx.y, &p.Rmatx.z, &p.Rmaty.x,
                        &p.Rmaty.y, &p.Rmaty.z, &p.Rmatz.x, &p.Rmatz.y, &p.Rmatz.z, &p.vv.x,
                        &p.vv.y, &p.vv.z, &p.ww.x, &p.ww.y, &p.ww.z);
    if (n != 18) ok = false;  // avalanche_hst.cu:245-249
  }
  std::fclose(f);
  return ok;
}

}  // namespace

int main(int argc, char **argv) {
  bool restart = false;
  const char *restart_bin = nullptr;
  int64_t max_steps = -1;
  const char *final_bin = nullptr;
  avf_options opt;
  avf_options_default(&opt);
  opt.verlet_skin = -1.0;  // chosen from the input below unless given
  opt.reorder_interval = 400;
  for (int i = 1; i < argc; i++) {
    if (!std::strcmp(argv[i], "--restart")) restart = true;
    else if (!std::strcmp(argv[i], "--restart-bin") && i + 1 < argc) restart_bin = argv[++i];
    else if (!std::strcmp(argv[i], "--max-steps") && i + 1 < argc) max_steps = std::atoll(argv[++i]);
    else if (!std::strcmp(argv[i], "--device") && i + 1 < argc) opt.device = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--final-bin") && i + 1 < argc) final_bin = argv[++i];
    else if (!std::strcmp(argv[i], "--skin") && i + 1 < argc) opt.verlet_skin = std::atof(argv[++i]);
    else if (!std::strcmp(argv[i], "--reorder") && i + 1 < argc)
      opt.reorder_interval = std::atoi(argv[++i]);
    else {
      std::fprintf(stderr, "usage: %s [--restart] [--restart-bin FILE] [--max-steps N] "
                           "[--final-bin FILE] [--device D] [--skin CM] [--reorder N] "
                           "< DataInput.dat\n", argv[0]);
      return 2;
    }
  }
  std::string text;
  char buf[4096];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof buf, stdin)) > 0) text.append(buf, got);

  avf_ctx *ctx = nullptr;
  avf_system sys;
  int64_t ntags = 0;
  int32_t print_flag = 0;
  int rc = avf_parse_datainput(text.c_str(), &sys, &ntags, &print_flag);
  if (rc) {
    std::fprintf(stderr, "malformed DataInput.dat (%d)\n", rc);
    return 1;
  }
  // a quarter of the largest sphere radius: the reference run (r = 0.18 cm, flows of a few
  // m/s) takes 143 s with it against 158 s at 0.02 cm (fewer list rebuilds)
  if (opt.verlet_skin < 0.0) opt.verlet_skin = 0.25 * sys.rmax;
  Outputs o;
  o.log = std::fopen("log_amd.dat", "w");
  std::fputs(text.c_str(), o.log);
  avf_geometry geo;
  avf_geometry_reference(&geo);
  avf_funnel fun;
  avf_derive_system(&sys, &fun, &geo);
  std::fprintf(o.log, "sinf: %lf, cosfi %lf, tanfi %lf\n", fun.sinfi, fun.cosfi, fun.tanfi);
  char msg[256];
  int errn = avf_error_check(&sys, msg, sizeof msg);
  if (errn) {  // avalanche_hst.cu:106-112
    std::fprintf(stderr, "%sError No. %d.\n", msg, errn);
    std::fprintf(o.log, "%sError No. %d.\n", msg, errn);
    return 1;
  }
  const double xfinal = (geo.lfunnel + sys.Lflume) * sys.cost;  // avalanche_hst.cu:116
  std::fprintf(o.log, "xfinal: %lf\n", xfinal);
  const int64_t nn = sys.npart;
  std::vector<avf_particle> pp((size_t)nn);
  std::vector<avf_double3> aa((size_t)nn), ga((size_t)nn);
  std::memset(pp.data(), 0, pp.size() * sizeof(avf_particle));
  avf_generate_particles(pp.data(), &sys, 2.6);
  avf_calculate_particle_parameters(pp.data(), &sys, &geo);
  sys.tot_mass = 0.0f;
  for (int64_t i = 0; i < nn; i++) sys.tot_mass += pp[i].mass;
  std::fprintf(o.log, "Total mass: %lf\n", sys.tot_mass);
  std::fprintf(o.log, "gamma_pp: %lf\ngamma_pw: %lf\nkappa_pp: %lf\nkappa_pw: %lf\n", sys.gamma_pp,
               sys.gamma_pw, sys.kappa_pp, sys.kappa_pw);
  std::fflush(o.log);

  int aval_start = 0;
  int64_t cc = 0, jj = 0;
  bool have_acc = false;
  if (restart_bin) {
    if (!read_bin(restart_bin, &sys, &aval_start, &cc, &jj, pp, aa, ga)) {
      std::fprintf(stderr, "Error reading input file\n");
      return 1;
    }
    have_acc = true;
  } else if (restart) {
    if (!read_text("last_config.dat", &sys, &aval_start, &cc, &jj, pp)) {
      std::fprintf(stderr, "Error reading input file\n");  // avalanche_hst.cu:247
      return 1;
    }
    avf_sphere_centres(pp.data(), nn, &geo);  // avalanche_hst.cu:265-293
  } else {
    if (avf_place_particles(pp.data(), &sys, &geo)) {
      std::printf("Particles don't fit!\n");  // place_particles.cu:58-62
      return 1;
    }
    avf_rng rng;
    avf_rng_init(&rng);
    avf_set_initial_velocity(pp.data(), &sys, &rng, 0);
  }
  if (aval_start && !restart_bin) {
    sys.side.x = xfinal;  // avalanche_hst.cu:254-263
    std::fprintf(o.log, "Hopper opened.\n");
  }
  const bool append = restart || restart_bin;
  o.xyz = std::fopen("snapshots_amd.xyz", append ? "a" : "w");
  o.vmean = std::fopen("mean_velocity_amd.csv", append ? "a" : "w");
  o.vel = std::fopen("velocities_amd.xyz", append ? "a" : "w");
  o.xcm = std::fopen("xcm_amd.csv", append ? "a" : "w");

  avf_instance_desc d;
  std::memset(&d, 0, sizeof d);
  d.sys = sys;
  d.geo = geo;
  d.fun = fun;
  d.ntags = ntags;
  d.phase = aval_start ? AVF_PHASE_AVALANCHE : AVF_PHASE_FILL;
  rc = avf_create(&ctx, &d, 1, &opt);
  if (rc) die("avf_create", nullptr, rc);
  CK(avf_upload(ctx, 0, pp.data(), have_acc ? aa.data() : nullptr, have_acc ? ga.data() : nullptr));
  CK(avf_set_phase(ctx, 0, d.phase, sys.time, sys.side.x, cc));
  CK(avf_regrid(ctx));
  if (!have_acc) CK(avf_initial_forces(ctx));

  // the stop tests read vmean as last set: before the first output step, the sum (not the
  // mean) of |v| over the clumps of the start state (avalanche_hst.cu:266-300), which only a
  // restart between output steps ever reads
  double vmean = 0.0;
  for (int64_t i = 0; i < nn; i++) {
    double vm2 = pp[i].vv.x * pp[i].vv.x;
    vm2 += pp[i].vv.y * pp[i].vv.y;
    vm2 += pp[i].vv.z * pp[i].vv.z;
    vmean += std::sqrt(vm2);
  }
  const auto t_start = std::chrono::steady_clock::now();
  int64_t total = 0;
  auto budget = [&](int64_t k) { return max_steps < 0 ? k : std::min<int64_t>(k, max_steps - total); };
  // One phase of the reference loop.  Steps run in chunks that end at the next output step
  // (cc % 20000 == 0), so every host-visible event happens between chunks.
  auto run_phase = [&](bool fill) {
    const double t_end = fill ? 0.0 : 2.0;
    while (sys.time < t_end) {
      if (max_steps >= 0 && total >= max_steps) return;
      // steps until the next output step (inclusive) or the end condition
      int64_t to_out = kOutputEvery - (cc % kOutputEvery);
      if (cc % kOutputEvery == 0) to_out = 1;
      int64_t k = 0;
      double t = sys.time;
      bool stop_after = false;
      while (k < to_out) {
        t += sys.dt;
        k++;
        if (!(t < t_end)) break;
        if (!fill && t > 0.2 && vmean < 0.1 * sys.v0) {  // avalanche_hst.cu:854-855
          stop_after = true;
          break;
        }
      }
      k = budget(k);
      if (k <= 0) return;
      CK(avf_step(ctx, k));
      for (int64_t i = 0; i < k; i++) sys.time += sys.dt;
      total += k;
      const int64_t last_cc = cc + k - 1;
      cc += k;
      if (last_cc % kOutputEvery == 0) {
        CK(avf_download(ctx, 0, pp.data(), aa.data(), ga.data()));
        std::printf(fill ? "Time = %lf s, %ld dt\n" : "Avalanche time = %lf s, %ld dt\n",
                    sys.time - sys.dt, (long)jj);
        avf_cell cell;
        CK(avf_get_cell(ctx, 0, &cell));
        std::printf("ncell: (%ld,%ld,%ld); ntot: %ld.\n", (long)cell.ncell.x, (long)cell.ncell.y,
                    (long)cell.ncell.z, (long)cell.ntot);
        avf_diag dg;
        CK(avf_get_diag(ctx, &dg));
        std::printf("steps %ld, candidate-list builds %ld, steps undone %ld, wall %.1f s\n",
                    (long)dg.steps, (long)dg.list_builds, (long)dg.steps_undone,
                    std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start)
                        .count());
        write_outputs(o, sys, pp, aval_start, last_cc, jj, print_flag, xfinal, geo.yfunnel, &vmean);
        write_bin("last_config.bin", sys, aval_start, last_cc + 1, jj + 1, pp, aa, ga);
        jj++;
        if (!fill && sys.time > 0.2 && vmean < 0.1 * sys.v0) return;
      }
      if (fill && vmean < 0.1 * sys.v0) return;  // avalanche_hst.cu:597-598
      if (stop_after) return;
    }
  };
  if (!aval_start) {
    std::printf("Filling hopper...\n");
    run_phase(true);
  }
  if (!aval_start && !(max_steps >= 0 && total >= max_steps)) {
    sys.side.x = xfinal;  // avalanche_hst.cu:606-613
    aval_start = 1;
    cc = 0;
    sys.time = 0.0;
    std::printf("Opening hopper... Avalanche starts!\n");
    CK(avf_set_phase(ctx, 0, AVF_PHASE_AVALANCHE, 0.0, xfinal, 0));
  }
  if (aval_start) run_phase(false);
  CK(avf_sync(ctx));
  if (final_bin) {  // lossless state at exit (restart with --restart-bin)
    CK(avf_download(ctx, 0, pp.data(), aa.data(), ga.data()));
    write_bin(final_bin, sys, aval_start, cc, jj, pp, aa, ga);
  }
  avf_destroy(ctx);
  std::fclose(o.vmean);
  std::fclose(o.vel);
  std::fclose(o.log);
  std::fclose(o.xyz);
  std::fclose(o.xcm);
  return 0;
}
