// Host check of contact_force_bwd against dual-number tangents of contact_force.
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <random>
#include "dem_window.h"
using namespace avf;

int main() {
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  int nbranch[4] = {0, 0, 0, 0};
  double worst = 0.0;
  for (int trial = 0; trial < 20000; trial++) {
    const bool pp = trial & 1;
    V<double> n = mv(U(rng), U(rng), U(rng));
    double nn = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
    n = mv(n.x / nn, n.y / nn, n.z / nn);
    const double scale = std::pow(10.0, -3.0 + 3.0 * (U(rng) + 1.0));
    V<double> dv = mv(scale * U(rng), scale * U(rng), scale * U(rng));
    if (trial % 4 >= 2) {  // nearly normal relative velocity: the stick branch
      const double a = scale * U(rng), tsc = std::pow(10.0, -14.0 + 6.0 * (U(rng) + 1.0) / 2.0);
      dv = mv(a * n.x + tsc * U(rng), a * n.y + tsc * U(rng), a * n.z + tsc * U(rng));
    }
    const double mass = 1e-4 * (1.5 + U(rng));
    const double comp = 1e-5 * (1.5 + U(rng));
    const double km = 0.5 * 1e6 * mass * (1.5 + U(rng)), gm = 0.5 * 30.0 * mass * (1.5 + U(rng));
    const double mu = 0.5 + 0.3 * U(rng);
    const double cns = (trial % 3 == 0) ? mu / 1e12 : mu * 1e-6 * (1.5 + U(rng));
    const double dt = 1e-6 * (1.5 + U(rng));
    V<double> f;
    if (!contact_force(dv, n, comp, km, gm, mu, cns, dt, mass, pp, f)) continue;
    // branch classification
    {
      double dn = dv.x * n.x + dv.y * n.y + dv.z * n.z;
      V<double> t = mv(dv.x - dn * n.x, dv.y - dn * n.y, dv.z - dn * n.z);
      double vt = std::sqrt(t.x * t.x + t.y * t.y + t.z * t.z);
      double normal = km * comp - gm * dn;
      double xi = dt * ((dv.x * t.x + dv.y * t.y + dv.z * t.z) / vt);
      bool reset = cns * normal < std::fabs(xi);
      if (reset) xi = 0.0;
      double fs = 1e12 * mass * xi + 2e6 * mass * xi / dt;
      bool slide = pp ? std::fabs(fs) > mu * normal : (std::fabs(fs) > mu * normal || xi == 0.0);
      nbranch[(slide ? 2 : 0) + (reset ? 1 : 0)]++;
    }
    V<double> fb = mv(U(rng), U(rng), U(rng));
    V<double> dvb = mv(0.0, 0.0, 0.0), nb = mv(0.0, 0.0, 0.0);
    double compb = 0, kmb = 0, gmb = 0, mub = 0;
    contact_force_bwd(dv, n, comp, km, gm, mu, cns, dt, mass, pp, fb, dvb, nb, compb, kmb, gmb, mub);
    // tangent direction
    double t_in[12];
    for (double &x : t_in) x = U(rng);
    V<Dual> dvD = mv(mkd(dv.x, t_in[0]), mkd(dv.y, t_in[1]), mkd(dv.z, t_in[2]));
    V<Dual> nD = mv(mkd(n.x, t_in[3]), mkd(n.y, t_in[4]), mkd(n.z, t_in[5]));
    Dual compD = mkd(comp, t_in[6] * 1e-5), kmD = mkd(km, t_in[7] * km), gmD = mkd(gm, t_in[8] * gm);
    Dual muD = mkd(mu, t_in[9]), cnsD = mkd(cns, 0.0);
    V<Dual> fD;
    contact_force(dvD, nD, compD, kmD, gmD, muD, cnsD, dt, mass, pp, fD);
    const double lhs = fb.x * fD.x.d + fb.y * fD.y.d + fb.z * fD.z.d;
    const double rhs = dvb.x * t_in[0] + dvb.y * t_in[1] + dvb.z * t_in[2] + nb.x * t_in[3] +
                       nb.y * t_in[4] + nb.z * t_in[5] + compb * t_in[6] * 1e-5 + kmb * t_in[7] * km +
                       gmb * t_in[8] * gm + mub * t_in[9];
    const double scl = std::fabs(lhs) + std::fabs(rhs) + 1e-300;
    const double rel = std::fabs(lhs - rhs) / scl;
    if (rel > worst) worst = rel;
    if (rel > 1e-9) {
      std::printf("trial %d pp %d rel %.3e lhs %.17g rhs %.17g\n", trial, pp, rel, lhs, rhs);
      return 1;
    }
  }
  std::printf("branches stick %d stick-reset %d slide %d slide-reset %d worst %.3e\n", nbranch[0],
              nbranch[1], nbranch[2], nbranch[3], worst);
  return 0;
}
