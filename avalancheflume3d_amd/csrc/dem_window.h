// dem_window.h -- differentiable DEM step over a window with a frozen contact list.
//
// The forward functions are templates over the scalar type: double for the forward and the
// recompute, Dual (value + one tangent) for the forward-mode check.  Every forward keeps
// the operation order of the reference's host ops (the same arithmetic as dem_math.h).  The
// reverse functions (double) are written by hand: each recomputes its forward
// intermediates and propagates adjoints line by line in reverse, following the branch the
// forward took (normal > 0, xi reset, stick / slide, |v_t| > 0).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace avf {

// ------------------------------------------------------------------------ dual numbers

// The double versions stay visible next to the Dual overloads declared below.
using ::cos;
using ::fabs;
using ::sin;
using ::sqrt;

struct Dual {
  double v, d;
};

__host__ __device__ __forceinline__ Dual mkd(double v, double d = 0.0) {
  Dual r;
  r.v = v;
  r.d = d;
  return r;
}
__host__ __device__ __forceinline__ Dual operator+(Dual a, Dual b) { return mkd(a.v + b.v, a.d + b.d); }
__host__ __device__ __forceinline__ Dual operator-(Dual a, Dual b) { return mkd(a.v - b.v, a.d - b.d); }
__host__ __device__ __forceinline__ Dual operator-(Dual a) { return mkd(-a.v, -a.d); }
__host__ __device__ __forceinline__ Dual operator*(Dual a, Dual b) {
  return mkd(a.v * b.v, a.d * b.v + a.v * b.d);
}
__host__ __device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return mkd(q, (a.d - q * b.d) / b.v);
}
__host__ __device__ __forceinline__ Dual operator+(Dual a, double b) { return mkd(a.v + b, a.d); }
__host__ __device__ __forceinline__ Dual operator+(double a, Dual b) { return mkd(a + b.v, b.d); }
__host__ __device__ __forceinline__ Dual operator-(Dual a, double b) { return mkd(a.v - b, a.d); }
__host__ __device__ __forceinline__ Dual operator-(double a, Dual b) { return mkd(a - b.v, -b.d); }
__host__ __device__ __forceinline__ Dual operator*(Dual a, double b) { return mkd(a.v * b, a.d * b); }
__host__ __device__ __forceinline__ Dual operator*(double a, Dual b) { return mkd(a * b.v, a * b.d); }
__host__ __device__ __forceinline__ Dual operator/(Dual a, double b) { return mkd(a.v / b, a.d / b); }
__host__ __device__ __forceinline__ Dual operator/(double a, Dual b) {
  const double q = a / b.v;
  return mkd(q, -q * b.d / b.v);
}
__host__ __device__ __forceinline__ Dual &operator+=(Dual &a, Dual b) { return a = a + b; }
__host__ __device__ __forceinline__ Dual &operator-=(Dual &a, Dual b) { return a = a - b; }
__host__ __device__ __forceinline__ Dual &operator/=(Dual &a, Dual b) { return a = a / b; }
__host__ __device__ __forceinline__ Dual sqrt(Dual a) {
  const double s = ::sqrt(a.v);
  return mkd(s, a.d / (2.0 * s));
}
__host__ __device__ __forceinline__ Dual sin(Dual a) { return mkd(::sin(a.v), ::cos(a.v) * a.d); }
__host__ __device__ __forceinline__ Dual cos(Dual a) { return mkd(::cos(a.v), -::sin(a.v) * a.d); }
__host__ __device__ __forceinline__ Dual fabs(Dual a) { return a.v < 0.0 ? -a : a; }

__host__ __device__ __forceinline__ double val(double a) { return a; }
__host__ __device__ __forceinline__ double val(Dual a) { return a.v; }
template <typename T>
__host__ __device__ __forceinline__ T lift(double a) {
  return T(a);
}
template <>
__host__ __device__ __forceinline__ Dual lift<Dual>(double a) {
  return mkd(a, 0.0);
}

template <typename T>
struct V {
  T x, y, z;
};
template <typename T>
__host__ __device__ __forceinline__ V<T> mv(T x, T y, T z) {
  V<T> r;
  r.x = x;
  r.y = y;
  r.z = z;
  return r;
}

// ------------------------------------------------------------------ per-instance data

// Differentiated contact-law parameters per instance, for the particle-particle (set 0)
// and particle-wall (set 1) laws: kappa (= 2 k_n / sphere mass), gamma (damping rate from
// the restitution coefficient), mu, and the tangential reset factor mu / 1e12; plus the
// tangent direction read by the dual-number forward.
enum WinSet { WS_PP = 0, WS_PW = 1 };
enum WinPar { WP_KAPPA = 0, WP_GAMMA = 1, WP_MU = 2, WP_FR = 3 };
struct WinParams {
  double v[2][4];
  double d[2][4];
};

// Adjoint accumulators of the contact-law parameters: pp (kappa, gamma, mu) and walls.
enum ThetaBar { TB_KPP, TB_GPP, TB_MPP, TB_KPW, TB_GPW, TB_MPW, TB_N };

// ------------------------------------------------------------ forward: contact laws ----

// Relative velocity decomposition, normal force and (if repulsive) the sphere force of one
// contact (force_part.cu:188-242; walls with pp_rule = false).  Returns whether the
// normal force is repulsive (nan_applies: the corner rule `normal <= 0.0` skip).
template <typename T>
__host__ __device__ __forceinline__ bool contact_force(V<T> dvvc, V<T> n, T comp, T km, T gm,
                                                       T mu, T cns_1, double dt, double mass_f,
                                                       bool pp_rule, V<T> &f) {
  T dvvcn_mag = dvvc.x * n.x + dvvc.y * n.y + dvvc.z * n.z;
  V<T> dvvcn = mv(dvvcn_mag * n.x, dvvcn_mag * n.y, dvvcn_mag * n.z);
  V<T> dvvct = mv(dvvc.x - dvvcn.x, dvvc.y - dvvcn.y, dvvc.z - dvvcn.z);
  T vt = sqrt(dvvct.x * dvvct.x + dvvct.y * dvvct.y + dvvct.z * dvvct.z);
  T normal = km * comp + gm * (-dvvcn_mag);
  if (!(val(normal) > 0.0)) return false;
  f = mv(-normal * n.x, -normal * n.y, -normal * n.z);
  if (!(val(vt) > 0.0)) return true;
  V<T> tv = mv(dvvct.x / vt, dvvct.y / vt, dvvct.z / vt);
  T sig = dvvc.x * tv.x + dvvc.y * tv.y + dvvc.z * tv.z;
  T xi = dt * sig;
  if (val(cns_1 * normal) < ::fabs(val(xi))) xi = lift<T>(0.0);
  T xi_dot = xi / dt;
  T ff_sdt = (1.0e12 * mass_f * xi + 2.0e6 * mass_f * xi_dot);
  bool slide = pp_rule ? (::fabs(val(ff_sdt)) > val(mu * normal))
                       : (::fabs(val(ff_sdt)) > val(mu * normal) || val(xi) == 0.0);
  if (slide) {
    T aux = 0.8 * mu * normal / vt;
    T t0 = mass_f / aux;
    aux /= (1.0 + (dt / t0) * (dt / t0));
    f.x += aux * dvvct.x;
    f.y += aux * dvvct.y;
    f.z += aux * dvvct.z;
  } else {
    if (val(dvvct.x) != 0.0) f.x += ff_sdt * dvvct.x / vt;
    if (val(dvvct.y) != 0.0) f.y += ff_sdt * dvvct.y / vt;
    if (val(dvvct.z) != 0.0) f.z += ff_sdt * dvvct.z / vt;
  }
  return true;
}

template <typename T>
__host__ __device__ __forceinline__ V<T> pvel(V<T> v, V<T> w, V<T> rrc, V<T> r) {
  return mv(v.x + w.y * (rrc.z - r.z) - w.z * (rrc.y - r.y),
            v.y + w.z * (rrc.x - r.x) - w.x * (rrc.z - r.z),
            v.z + w.x * (rrc.y - r.y) - w.y * (rrc.x - r.x));
}

template <typename T>
__host__ __device__ __forceinline__ void add_ft(V<T> &ff, V<T> &tq, V<T> rrc, V<T> r, V<T> f) {
  tq.x += (rrc.y - r.y) * f.z - (rrc.z - r.z) * f.y;
  tq.y += (rrc.z - r.z) * f.x - (rrc.x - r.x) * f.z;
  tq.z += (rrc.x - r.x) * f.y - (rrc.y - r.y) * f.x;
  ff.x += f.x;
  ff.y += f.y;
  ff.z += f.z;
}

// ------------------------------------------------------------- reverse: contact law ----

// Adjoints of the inputs of contact_force given fbar (the force adjoint).  Recomputes the
// forward; branches follow the forward.  Outputs accumulate (+=).
__host__ __device__ inline void contact_force_bwd(V<double> dvvc, V<double> n, double comp,
                                                  double km, double gm, double mu, double cns_1,
                                                  double dt, double mass_f, bool pp_rule,
                                                  V<double> fb, V<double> &dvvc_b,
                                                  V<double> &n_b, double &comp_b, double &km_b,
                                                  double &gm_b, double &mu_b) {
  const double dn = dvvc.x * n.x + dvvc.y * n.y + dvvc.z * n.z;
  const V<double> dvvct = mv(dvvc.x - dn * n.x, dvvc.y - dn * n.y, dvvc.z - dn * n.z);
  const double vt = ::sqrt(dvvct.x * dvvct.x + dvvct.y * dvvct.y + dvvct.z * dvvct.z);
  const double normal = km * comp + gm * (-dn);
  if (!(normal > 0.0)) return;
  double normal_b = 0.0, vt_b = 0.0, sig_b = 0.0;
  V<double> dvvct_b = mv(0.0, 0.0, 0.0), tv_b = mv(0.0, 0.0, 0.0);
  // f = -normal n + friction
  normal_b += -(fb.x * n.x + fb.y * n.y + fb.z * n.z);
  n_b.x += -normal * fb.x;
  n_b.y += -normal * fb.y;
  n_b.z += -normal * fb.z;
  if (vt > 0.0) {
    const V<double> tv = mv(dvvct.x / vt, dvvct.y / vt, dvvct.z / vt);
    const double sig = dvvc.x * tv.x + dvvc.y * tv.y + dvvc.z * tv.z;
    double xi = dt * sig;
    const bool reset = cns_1 * normal < ::fabs(xi);
    if (reset) xi = 0.0;
    const double xi_dot = xi / dt;
    const double ff_sdt = (1.0e12 * mass_f * xi + 2.0e6 * mass_f * xi_dot);
    const bool slide = pp_rule ? (::fabs(ff_sdt) > mu * normal)
                               : (::fabs(ff_sdt) > mu * normal || xi == 0.0);
    const double fdv = fb.x * dvvct.x + fb.y * dvvct.y + fb.z * dvvct.z;
    if (slide) {
      // aux2 = aux / (1 + q aux^2), q = (dt / mass_f)^2, aux = 0.8 mu N / vt
      const double aux = 0.8 * mu * normal / vt;
      const double q = (dt / mass_f) * (dt / mass_f);
      const double den = 1.0 + q * aux * aux;
      const double aux2 = aux / den;
      dvvct_b.x += aux2 * fb.x;
      dvvct_b.y += aux2 * fb.y;
      dvvct_b.z += aux2 * fb.z;
      const double aux_b = fdv * (1.0 - q * aux * aux) / (den * den);
      mu_b += aux_b * 0.8 * normal / vt;
      normal_b += aux_b * 0.8 * mu / vt;
      vt_b += -aux_b * aux / vt;
    } else {
      // f += ff_sdt dvvct / vt (components with dvvct == 0 contribute 0 either way)
      dvvct_b.x += ff_sdt / vt * fb.x;
      dvvct_b.y += ff_sdt / vt * fb.y;
      dvvct_b.z += ff_sdt / vt * fb.z;
      vt_b += -ff_sdt * fdv / (vt * vt);
      if (!reset) {
        const double sdt_b = fdv / vt;
        const double xi_b = sdt_b * (1.0e12 * mass_f + 2.0e6 * mass_f / dt);
        sig_b += xi_b * dt;
      }
    }
    // sig = dvvc . tv
    dvvc_b.x += sig_b * tv.x;
    dvvc_b.y += sig_b * tv.y;
    dvvc_b.z += sig_b * tv.z;
    tv_b.x += sig_b * dvvc.x;
    tv_b.y += sig_b * dvvc.y;
    tv_b.z += sig_b * dvvc.z;
    // tv = dvvct / vt
    const double tvd = tv_b.x * dvvct.x + tv_b.y * dvvct.y + tv_b.z * dvvct.z;
    dvvct_b.x += tv_b.x / vt;
    dvvct_b.y += tv_b.y / vt;
    dvvct_b.z += tv_b.z / vt;
    vt_b += -tvd / (vt * vt);
    // vt = |dvvct|
    dvvct_b.x += vt_b * dvvct.x / vt;
    dvvct_b.y += vt_b * dvvct.y / vt;
    dvvct_b.z += vt_b * dvvct.z / vt;
  }
  // normal = km comp - gm dn
  km_b += normal_b * comp;
  comp_b += normal_b * km;
  gm_b += -normal_b * dn;
  double dn_b = -normal_b * gm;
  // dvvct = dvvc - dn n
  dvvc_b.x += dvvct_b.x;
  dvvc_b.y += dvvct_b.y;
  dvvc_b.z += dvvct_b.z;
  dn_b += -(dvvct_b.x * n.x + dvvct_b.y * n.y + dvvct_b.z * n.z);
  n_b.x += -dn * dvvct_b.x;
  n_b.y += -dn * dvvct_b.y;
  n_b.z += -dn * dvvct_b.z;
  // dn = dvvc . n
  dvvc_b.x += dn_b * n.x;
  dvvc_b.y += dn_b * n.y;
  dvvc_b.z += dn_b * n.z;
  n_b.x += dn_b * dvvc.x;
  n_b.y += dn_b * dvvc.y;
  n_b.z += dn_b * dvvc.z;
}

// y = w x e: adjoints w_b += e x y_b, e_b += y_b x w.
__host__ __device__ __forceinline__ void cross_bwd(V<double> w, V<double> e, V<double> yb,
                                                   V<double> &wb, V<double> &eb) {
  wb.x += e.y * yb.z - e.z * yb.y;
  wb.y += e.z * yb.x - e.x * yb.z;
  wb.z += e.x * yb.y - e.y * yb.x;
  eb.x += yb.y * w.z - yb.z * w.y;
  eb.y += yb.z * w.x - yb.x * w.z;
  eb.z += yb.x * w.y - yb.y * w.x;
}

// Adjoint of "vc = v + w x (rrc - r)" given vc_b.
__host__ __device__ __forceinline__ void pvel_bwd(V<double> w, V<double> rrc, V<double> r,
                                                  V<double> vcb, V<double> &vb, V<double> &wb,
                                                  V<double> &rrcb, V<double> &rb) {
  vb.x += vcb.x;
  vb.y += vcb.y;
  vb.z += vcb.z;
  V<double> e = mv(rrc.x - r.x, rrc.y - r.y, rrc.z - r.z), eb = mv(0.0, 0.0, 0.0);
  cross_bwd(w, e, vcb, wb, eb);
  rrcb.x += eb.x;
  rrcb.y += eb.y;
  rrcb.z += eb.z;
  rb.x -= eb.x;
  rb.y -= eb.y;
  rb.z -= eb.z;
}

// Adjoint of "tq += (rrc - r) x f; ff += f" given ff_b, tq_b: returns f_b, accumulates
// rrc_b and r_b.
__host__ __device__ __forceinline__ V<double> add_ft_bwd(V<double> rrc, V<double> r, V<double> f,
                                                         V<double> ffb, V<double> tqb,
                                                         V<double> &rrcb, V<double> &rb) {
  V<double> d = mv(rrc.x - r.x, rrc.y - r.y, rrc.z - r.z);
  // tau = d x f: d_b = f x tau_b, f_b = tau_b x d
  V<double> db = mv(f.y * tqb.z - f.z * tqb.y, f.z * tqb.x - f.x * tqb.z, f.x * tqb.y - f.y * tqb.x);
  V<double> fb = mv(ffb.x + tqb.y * d.z - tqb.z * d.y, ffb.y + tqb.z * d.x - tqb.x * d.z,
                    ffb.z + tqb.x * d.y - tqb.y * d.x);
  rrcb.x += db.x;
  rrcb.y += db.y;
  rrcb.z += db.z;
  rb.x -= db.x;
  rb.y -= db.y;
  rb.z -= db.z;
  return fb;
}

}  // namespace avf
