// dem_window.hip -- forward, tangent and adjoint kernels of the differentiable window.
//
// Over a window the contact list is frozen: every directed sphere pair (l <- m) closer than
// sep + skin at the window start, stored per clump in the order force_part.cu accumulates
// (pairing sph_id major, then clump cell offset, then reference index).  Each step is
//   k_win_predict   init_Verlet + sphere centres -> stage (x', v*, w*, R', spheres)
//   k_win_force     boundary modules + listed contacts + get_accel + finish_Verlet
// and the adjoint of one step is
//   k_win_predict   (recomputed stage of the step from its stored pre-step state)
//   k_win_adj_force reverse of finish, accel, walls and the listed contacts of clump l;
//                   the adjoint owed to the neighbour m of each entry goes to ebuf
//   k_win_adj_gather  per clump m: its own stage adjoint + the ebuf entries that name it
//                   (transposed list, fixed order), then the reverse of the sphere centres
//                   and of init_Verlet.
// All adjoint sums are sequential in a fixed order: no float atomics.
#include "dem_window.h"

namespace avf {

constexpr int kWinMaxNeigh = 48;   // touching-or-near clumps per clump at the window start
constexpr int kWinMaxList = 128;   // listed sphere pairs per clump
constexpr int kWinMaxIn = 160;     // listed entries naming a clump
constexpr int kStageW = 30;        // x'(3) v*(3) w*(3) R'(9) spheres(12)
enum WinStage { W_X = 0, W_V = 3, W_W = 6, W_R = 9, W_S = 18 };
constexpr int kEbuf = 12;          // adjoint owed to m: x'(3), v*(3), w*(3), its sphere (3)

template <typename T>
__device__ __forceinline__ V<T> ld3(const T *p) {
  return mv(p[0], p[1], p[2]);
}
template <typename T>
__device__ __forceinline__ void st3(T *p, V<T> v) {
  p[0] = v.x;
  p[1] = v.y;
  p[2] = v.z;
}

template <typename T>
__device__ __forceinline__ T wpar(const WinParams &w, int k);
template <>
__device__ __forceinline__ double wpar<double>(const WinParams &w, int k) {
  return k == 0 ? w.kappa : k == 1 ? w.gamma : k == 2 ? w.mu : w.fr;
}
template <>
__device__ __forceinline__ Dual wpar<Dual>(const WinParams &w, int k) {
  return k == 0 ? mkd(w.kappa, w.dkappa)
                : k == 1 ? mkd(w.gamma, w.dgamma) : k == 2 ? mkd(w.mu, w.dmu) : mkd(w.fr, w.dfr);
}

// ------------------------------------------------------------- list construction ----

// Frozen list of one clump from the binned current state: sphere pairs within sep + skin,
// grouped by touching clump (sorted by the reference key) and emitted pairing-major.
__global__ __launch_bounds__(64) void k_win_list(
    const double *__restrict__ st, const double *__restrict__ sta,
    const double *__restrict__ stage, const InstParams *__restrict__ P,
    const int32_t *__restrict__ inst_of, const uint32_t *__restrict__ starts,
    const double4 *__restrict__ spos, const uint32_t *__restrict__ sid, double skin,
    int32_t *__restrict__ list, int32_t *__restrict__ cnt_out, DevDiag *__restrict__ diag,
    int64_t n, int64_t cap) {
  __shared__ int32_t nm[kWinMaxNeigh][64];
  __shared__ uint32_t nk[kWinMaxNeigh][64];
  __shared__ uint16_t nmask[kWinMaxNeigh][64];
  const int lane = threadIdx.x;
  const int64_t c = (int64_t)blockIdx.x * 64 + lane;
  if (c >= n) return;
  const InstParams &p = P[inst_of[c]];
  const double *gl = stage + c * kStageDoubles;
  const d3 rr = make_d3(gl[G_RR], gl[G_RR + 1], gl[G_RR + 2]);
  const double rad = gl[G_RAD];
  int cnt = 0;
  bool overflow = false;
  int64_t li = (int64_t)((rr.x - p.rr_min[0]) / p.cell_side[0]);
  if (li == p.ncell[0]) li--;
  int64_t lj = (int64_t)((rr.y - p.rr_min[1]) / p.cell_side[1]);
  if (lj == p.ncell[1]) lj--;
  int64_t lk = (int64_t)((rr.z - p.rr_min[2]) / p.cell_side[2]);
  if (lk == p.ncell[2]) lk--;
  const int64_t nxy = p.ncell[0] * p.ncell[1];
  const int nx = p.sg_n[0], ny = p.sg_n[1], nz = p.sg_n[2];
  // sphere cells must cover sep + skin: widen the sweep by the skin in cell units
  const int reach = 1 + (int)ceil(skin * p.sg_inv_h);
  for (int s = 0; s < 4; s++) {
    const d3 q = make_d3(gl[G_SPH + 3 * s], gl[G_SPH + 3 * s + 1], gl[G_SPH + 3 * s + 2]);
    const int ix = grid_index(q.x, p.sg_lo[0], p.sg_inv_h, nx);
    const int iy = grid_index(q.y, p.sg_lo[1], p.sg_inv_h, ny);
    const int iz = grid_index(q.z, p.sg_lo[2], p.sg_inv_h, nz);
    const int x0 = ix - reach > 0 ? ix - reach : 0, x1 = ix + reach < nx ? ix + reach : nx - 1;
    for (int z = iz - reach; z <= iz + reach; z++) {
      if (z < 0 || z >= nz) continue;
      for (int y = iy - reach; y <= iy + reach; y++) {
        if (y < 0 || y >= ny) continue;
        const int64_t row = p.sg_base + (int64_t)nx * (y + (int64_t)ny * z);
        for (uint32_t j = starts[row + x0]; j < starts[row + x1 + 1]; j++) {
          const uint32_t id = sid[j];
          const int64_t cm = (int64_t)(id >> 2);
          if (cm == c) continue;
          const double4 o = spos[j];
          const double lim = rad + o.w + skin;
          const double dx = o.x - q.x, dy = o.y - q.y, dz = o.z - q.z;
          if (!(dx * dx + dy * dy + dz * dz < lim * lim)) continue;
          const uint32_t bit = 1u << (((id & 3u) << 2) | (uint32_t)s);
          int e = 0;
          for (; e < cnt; e++)
            if (nm[e][lane] == (int32_t)cm) break;
          if (e < cnt) {
            nmask[e][lane] = (uint16_t)(nmask[e][lane] | bit);
            continue;
          }
          const double *g = stage + cm * kStageDoubles;
          const int64_t L = ref_cell_locate(p, make_d3(g[G_RR], g[G_RR + 1], g[G_RR + 2]));
          if (L < 0) continue;
          const int64_t ak = L / nxy, rem = L - ak * nxy;
          const int64_t aj = rem / p.ncell[0], ai = rem - aj * p.ncell[0];
          const int64_t di = ai - li, dj = aj - lj, dk = ak - lk;
          // beyond the reference's 27 clump cells: keep it after them (order key 27)
          const bool near = di >= -1 && di <= 1 && dj >= -1 && dj <= 1 && dk >= -1 && dk <= 1;
          const uint32_t off = near ? (uint32_t)((di + 1) * 9 + (dj + 1) * 3 + (dk + 1)) : 27u;
          const uint32_t key = (off << 27) | (uint32_t)__double_as_longlong(g[G_ID]);
          if (cnt == kWinMaxNeigh) {
            overflow = true;
            continue;
          }
          int pos = cnt;
          while (pos > 0 && nk[pos - 1][lane] > key) {
            nk[pos][lane] = nk[pos - 1][lane];
            nm[pos][lane] = nm[pos - 1][lane];
            nmask[pos][lane] = nmask[pos - 1][lane];
            pos--;
          }
          nk[pos][lane] = key;
          nm[pos][lane] = (int32_t)cm;
          nmask[pos][lane] = (uint16_t)bit;
          cnt++;
        }
      }
    }
  }
  int out = 0;
  for (int pr = 0; pr < 16; pr++)
    for (int e = 0; e < cnt; e++)
      if ((nmask[e][lane] >> pr) & 1) {
        if (out < kWinMaxList)
          list[c * kWinMaxList + out] = (nm[e][lane] << 4) | pr;
        else
          overflow = true;
        out++;
      }
  cnt_out[c] = out < kWinMaxList ? out : kWinMaxList;
  if (overflow) atomicAdd(&diag->neigh_overflow, 1ull);
  atomicMax(&diag->max_neighbours, (unsigned long long)out);
}

// Transposed list: for every clump m, the entries (l * kWinMaxList + e) that name it,
// sorted ascending so the adjoint gather sums them in a fixed order.
__global__ void k_win_in_fill(const int32_t *__restrict__ list, const int32_t *__restrict__ cnt,
                              int32_t *__restrict__ in, int32_t *__restrict__ icnt,
                              DevDiag *__restrict__ diag, int64_t n) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  for (int e = 0; e < cnt[l]; e++) {
    const int32_t m = list[l * kWinMaxList + e] >> 4;
    const int slot = atomicAdd(&icnt[m], 1);
    if (slot < kWinMaxIn)
      in[(int64_t)m * kWinMaxIn + slot] = (int32_t)(l * kWinMaxList + e);
    else
      atomicAdd(&diag->neigh_overflow, 1ull);
  }
}

__global__ void k_win_in_sort(int32_t *__restrict__ in, int32_t *__restrict__ icnt, int64_t n) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  int k = icnt[m];
  if (k > kWinMaxIn) k = icnt[m] = kWinMaxIn;
  int32_t *a = in + m * kWinMaxIn;
  for (int i = 1; i < k; i++) {
    const int32_t v = a[i];
    int j = i - 1;
    while (j >= 0 && a[j] > v) {
      a[j + 1] = a[j];
      j--;
    }
    a[j + 1] = v;
  }
}

// ------------------------------------------------------------------- forward step ----

// init_Verlet.cu:46-148 for one clump: predicted x', v*, w*, R' and the sphere centres.
template <typename T>
__device__ __forceinline__ void win_predict_one(const T *S, double dt, double hdt, double pos, T *g) {
  V<T> rr = mv(S[S_RX], S[S_RY], S[S_RZ]), vv = mv(S[S_VX], S[S_VY], S[S_VZ]);
  V<T> ww = mv(S[S_WX], S[S_WY], S[S_WZ]), aa = mv(S[S_AX], S[S_AY], S[S_AZ]);
  V<T> ang = mv(S[S_GX], S[S_GY], S[S_GZ]);
  T R[9];
  for (int k = 0; k < 9; k++) R[k] = S[S_R00 + k];
  rr.x += dt * (vv.x + hdt * aa.x);
  rr.y += dt * (vv.y + hdt * aa.y);
  rr.z += dt * (vv.z + hdt * aa.z);
  vv.x += dt * aa.x;
  vv.y += dt * aa.y;
  vv.z += dt * aa.z;
  V<T> dO = mv(dt * (ww.x + hdt * ang.x), dt * (ww.y + hdt * ang.y), dt * (ww.z + hdt * ang.z));
  T mag = sqrt(dO.x * dO.x + dO.y * dO.y + dO.z * dO.z);
  T dR[9];
  if (val(mag) > 0.0) {
    V<T> nn = mv(dO.x / mag, dO.y / mag, dO.z / mag);
    T cdO = cos(mag), sdO = sin(mag), om = 1.0 - cdO;
    dR[0] = om * nn.x * nn.x + cdO;
    dR[1] = om * nn.x * nn.y - sdO * nn.z;
    dR[2] = om * nn.x * nn.z + sdO * nn.y;
    dR[3] = om * nn.y * nn.x + sdO * nn.z;
    dR[4] = om * nn.y * nn.y + cdO;
    dR[5] = om * nn.y * nn.z - sdO * nn.x;
    dR[6] = om * nn.z * nn.x - sdO * nn.y;
    dR[7] = om * nn.z * nn.y + sdO * nn.x;
    dR[8] = om * nn.z * nn.z + cdO;
  } else {
    // the smooth limit I + [dOmega]x (equal to I in value, carries the tangent)
    dR[0] = lift<T>(1.0);
    dR[1] = -dO.z;
    dR[2] = dO.y;
    dR[3] = dO.z;
    dR[4] = lift<T>(1.0);
    dR[5] = -dO.x;
    dR[6] = -dO.y;
    dR[7] = dO.x;
    dR[8] = lift<T>(1.0);
  }
  T Tm[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      Tm[3 * i + j] = dR[3 * i] * R[j] + dR[3 * i + 1] * R[3 + j] + dR[3 * i + 2] * R[6 + j];
  for (int j = 0; j < 3; j++) {
    T norm = sqrt(Tm[j] * Tm[j] + Tm[3 + j] * Tm[3 + j] + Tm[6 + j] * Tm[6 + j]);
    R[j] = Tm[j] / norm;
    R[3 + j] = Tm[3 + j] / norm;
    R[6 + j] = Tm[6 + j] / norm;
  }
  ww.x += dt * ang.x;
  ww.y += dt * ang.y;
  ww.z += dt * ang.z;
  st3(g + W_X, rr);
  st3(g + W_V, vv);
  st3(g + W_W, ww);
  for (int k = 0; k < 9; k++) g[W_R + k] = R[k];
  const double sg[4][3] = {{1, 1, 1}, {1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}};
  for (int s = 0; s < 4; s++) {
    g[W_S + 3 * s + 0] = rr.x + pos * (sg[s][0] * R[0] + sg[s][1] * R[1] + sg[s][2] * R[2]);
    g[W_S + 3 * s + 1] = rr.y + pos * (sg[s][0] * R[3] + sg[s][1] * R[4] + sg[s][2] * R[5]);
    g[W_S + 3 * s + 2] = rr.z + pos * (sg[s][0] * R[6] + sg[s][1] * R[7] + sg[s][2] * R[8]);
  }
}

// State S (field-major SoA of T) -> stage (AoS of T, kStageW per clump).
template <typename T>
__global__ void k_win_predict(const T *__restrict__ st, const double *__restrict__ sta,
                              const InstParams *__restrict__ P, const int32_t *__restrict__ inst_of,
                              T *__restrict__ stage, int64_t n, int64_t cap) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const InstParams &p = P[inst_of[c]];
  T S[S_NFIELDS];
  for (int f = 0; f < S_NFIELDS; f++) S[f] = st[f * cap + c];
  win_predict_one(S, p.dt, p.hdt, sta[C_RAD0 * cap + c] / p.sphere_div, stage + c * kStageW);
}

// Wall contact of sphere sp (lab frame) given its contact point, normal and compression.
template <typename T>
__device__ __forceinline__ void win_wall(V<T> rr, V<T> vv, V<T> ww, V<T> rrc, V<T> n, T comp,
                                         T km, T gm, T mu, T cns, double dt, double mass,
                                         V<T> &ff, V<T> &tq) {
  V<T> vc = pvel(vv, ww, rrc, rr);
  V<T> f;
  if (contact_force(mv(-vc.x, -vc.y, -vc.z), n, comp, km, gm, mu, cns, dt, mass, false, f))
    add_ft(ff, tq, rrc, rr, f);
}

// Boundary modules of the window (avalanche phase): hopper x / y walls, flume side walls,
// flume floor, deposit floor, in the reference order (hop, walls, [pp], flume, dep).
template <typename T>
__device__ __forceinline__ void win_walls_pre(const InstParams &p, int mods, V<T> rr, V<T> vv,
                                              V<T> ww, const T *sph, double ri, double mass,
                                              const WinParams &wp, V<T> &ff, V<T> &tq) {
  const T kap = wpar<T>(wp, 0), gam = wpar<T>(wp, 1), mu = wpar<T>(wp, 2), cns = wpar<T>(wp, 3);
  const T km = 0.5 * kap * mass, gm = 0.5 * gam * mass;
  if (mods & AVF_MOD_HOP) {
    for (int nw = 0; nw < 2; nw++)
      for (int ns = 0; ns < 4; ns++) {
        V<T> r = ld3(sph + 3 * ns);
        if (val(r.z) < p.hop_height || val(r.x) > p.hop_width) continue;
        int touch = 0;
        V<T> rrc, nv;
        T comp;
        if (nw == 0) {
          T dist = r.x - 0.0;
          if (val(dist) <= ri) {
            touch = 1;
            rrc = mv(lift<T>(0.0), r.y, r.z);
            nv = mv(lift<T>(-1.0), lift<T>(0.0), lift<T>(0.0));
            comp = ri - dist;
          }
          dist = p.hop_width - r.x;
          if (val(dist) <= ri) {
            touch = 2;
            rrc = mv(lift<T>(p.hop_width), r.y, r.z);
            nv = mv(lift<T>(1.0), lift<T>(0.0), lift<T>(0.0));
            comp = ri - dist;
          }
        } else {
          T dist = r.y - p.yhop_l;
          if (val(dist) <= ri) {
            touch = 3;
            rrc = mv(r.x, lift<T>(p.yhop_l), r.z);
            nv = mv(lift<T>(0.0), lift<T>(-1.0), lift<T>(0.0));
            comp = ri - dist;
          }
          dist = p.yhop_r - r.y;
          if (val(dist) <= ri) {
            touch = 4;
            rrc = mv(r.x, lift<T>(p.yhop_r), r.z);
            nv = mv(lift<T>(0.0), lift<T>(1.0), lift<T>(0.0));
            comp = ri - dist;
          }
        }
        if (touch) win_wall(rr, vv, ww, rrc, nv, comp, km, gm, mu, cns, p.dt, mass, ff, tq);
      }
  }
  if (mods & AVF_MOD_WALLS) {
    const double yl = p.dy2, yr = p.side_y + p.dy2;
    for (int ns = 0; ns < 4; ns++) {
      V<T> r = ld3(sph + 3 * ns);
      const double qx = val(r.x) * p.cost - val(r.z) * p.sint;
      if (qx <= p.lfunnel) continue;
      int touch = 0;
      V<T> rrc, nv;
      T comp;
      T dist = r.y - yl;
      if (val(dist) <= ri) {
        touch = 1;
        rrc = mv(r.x, lift<T>(yl), r.z);
        nv = mv(lift<T>(0.0), lift<T>(-1.0), lift<T>(0.0));
        comp = ri - dist;
      }
      dist = yr - r.y;
      if (val(dist) <= ri) {
        touch = 2;
        rrc = mv(r.x, lift<T>(yr), r.z);
        nv = mv(lift<T>(0.0), lift<T>(1.0), lift<T>(0.0));
        comp = ri - dist;
      }
      if (touch) win_wall(rr, vv, ww, rrc, nv, comp, km, gm, mu, cns, p.dt, mass, ff, tq);
    }
  }
}

template <typename T>
__device__ __forceinline__ void win_walls_post(const InstParams &p, int mods, V<T> rr, V<T> vv,
                                               V<T> ww, const T *sph, double ri, double mass,
                                               const WinParams &wp, V<T> &ff, V<T> &tq) {
  const T kap = wpar<T>(wp, 0), gam = wpar<T>(wp, 1), mu = wpar<T>(wp, 2), cns = wpar<T>(wp, 3);
  const T km = 0.5 * kap * mass, gm = 0.5 * gam * mass;
  if (mods & AVF_MOD_FLUME) {
    const double sint = p.sint, cost = p.cost, tant = sint / cost;
    for (int ns = 0; ns < 4; ns++) {
      V<T> r = ld3(sph + 3 * ns);
      T rect = -tant * r.x - r.z;
      T dist = fabs(cost * rect);
      if (!(val(dist) <= ri)) continue;
      V<T> nv = mv(lift<T>(-sint), lift<T>(0.0), lift<T>(-cost));
      V<T> rrc = mv(r.x - dist * sint, r.y, r.z - dist * cost);
      win_wall(rr, vv, ww, rrc, nv, ri - dist, km, gm, mu, cns, p.dt, mass, ff, tq);
    }
  }
  if (mods & AVF_MOD_DEP) {
    const double sind = p.dep_sin, cosd = p.dep_cos, tand = sind / cosd;
    for (int ns = 0; ns < 4; ns++) {
      V<T> r = ld3(sph + 3 * ns);
      T rect = tand * (r.x - p.dep_x0 - p.dep_x1) - (r.z - p.dep_z0);
      T dist = cosd * (-rect);
      if (!(val(dist) <= ri)) continue;
      V<T> nv = mv(lift<T>(sind), lift<T>(0.0), lift<T>(-cosd));
      T rx = r.x + dist * sind;
      V<T> rrc = mv(rx, r.y, tand * (rx - p.dep_x0 - p.dep_x1) + p.dep_z0);
      win_wall(rr, vv, ww, rrc, nv, ri - dist, km, gm, mu, cns, p.dt, mass, ff, tq);
    }
  }
}

// Listed contacts, get_accel and finish_Verlet for one clump: stage -> new state.
template <typename T>
__global__ __launch_bounds__(64) void k_win_force(
    const T *__restrict__ st, T *__restrict__ st_out, const double *__restrict__ sta,
    const T *__restrict__ stage, const InstParams *__restrict__ P,
    const WinParams *__restrict__ WP, const int32_t *__restrict__ inst_of,
    const int32_t *__restrict__ list, const int32_t *__restrict__ cnt, int64_t n, int64_t cap) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int inst = inst_of[c];
  const InstParams &p = P[inst];
  const WinParams &wp = WP[inst];
  const T *g = stage + c * kStageW;
  const V<T> rr = ld3(g + W_X);
  V<T> vv = ld3(g + W_V), ww = ld3(g + W_W);
  const double rad = sta[C_RAD * cap + c], rho = sta[C_RHO * cap + c];
  const double massw = kSphFact * rad * rad * rad * rho;
  const int mods = p.modules_aval;
  V<T> ff = mv(lift<T>(0.0), lift<T>(0.0), lift<T>(0.0)), tq = ff;
  win_walls_pre(p, mods, rr, vv, ww, g + W_S, rad, massw, wp, ff, tq);
  if (mods & AVF_MOD_PP) {
    const double massl = kSphFact * rho * rad * rad * rad;
    const T kap = wpar<T>(wp, 0), gam = wpar<T>(wp, 1), mu = wpar<T>(wp, 2), cns = wpar<T>(wp, 3);
    const int k = cnt[c];
    for (int e = 0; e < k; e++) {
      const int32_t ent = list[c * kWinMaxList + e];
      const int64_t m = ent >> 4;
      const int pr = ent & 15, sdiv = pr >> 2, smod = pr & 3;
      const T *gm_ = stage + m * kStageW;
      const double rad2 = sta[C_RAD * cap + m], rho2 = sta[C_RHO * cap + m];
      const V<T> r1 = ld3(g + W_S + 3 * smod), r2 = ld3(gm_ + W_S + 3 * sdiv);
      const double sep = rad + rad2;
      V<T> drr = mv(r2.x - r1.x, r2.y - r1.y, r2.z - r1.z);
      T dist2 = drr.x * drr.x + drr.y * drr.y + drr.z * drr.z;
      if (!(val(dist2) < sep * sep)) continue;
      const double massm = kSphFact * rho2 * rad2 * rad2 * rad2;
      const double meff = 0.5 * (massm + massl);
      const T km = 0.5 * kap * meff, gmm = 0.5 * gam * meff;
      T dist = sqrt(dist2);
      T comp = sep - dist;
      V<T> nv = mv(drr.x / dist, drr.y / dist, drr.z / dist);
      V<T> rrc = mv(r1.x + rad * nv.x, r1.y + rad * nv.y, r1.z + rad * nv.z);
      const V<T> xm = ld3(gm_ + W_X), vm = ld3(gm_ + W_V), wm = ld3(gm_ + W_W);
      V<T> vl = pvel(vv, ww, rrc, rr), vmc = pvel(vm, wm, rrc, xm);
      V<T> f;
      if (contact_force(mv(vmc.x - vl.x, vmc.y - vl.y, vmc.z - vl.z), nv, comp, km, gmm, mu, cns,
                        p.dt, massl, true, f))
        add_ft(ff, tq, rrc, rr, f);
    }
  }
  win_walls_post(p, mods, rr, vv, ww, g + W_S, rad, massw, wp, ff, tq);
  // get_accel.cu:37-71 (isotropic inertia: the gyroscopic terms vanish)
  const double mass = sta[C_MASS * cap + c], I = sta[C_IX * cap + c];
  T R[9];
  for (int k = 0; k < 9; k++) R[k] = g[W_R + k];
  const double cd = p.drag * rad * rad / mass;
  T vmag = sqrt(vv.x * vv.x + vv.y * vv.y + vv.z * vv.z);
  T adrag = cd * vmag;
  V<T> af = mv(ff.x / mass - adrag * vv.x, ff.y / mass - adrag * vv.y,
               ff.z / mass - adrag * vv.z - p.grav);
  V<T> tb = mv(tq.x * R[0] + tq.y * R[3] + tq.z * R[6], tq.x * R[1] + tq.y * R[4] + tq.z * R[7],
               tq.x * R[2] + tq.y * R[5] + tq.z * R[8]);
  V<T> ab = mv(tb.x / I, tb.y / I, tb.z / I);
  V<T> gf = mv(ab.x * R[0] + ab.y * R[1] + ab.z * R[2], ab.x * R[3] + ab.y * R[4] + ab.z * R[5],
               ab.x * R[6] + ab.y * R[7] + ab.z * R[8]);
  // finish_Verlet.cu:43-48
  const double hdt = p.hdt;
  const V<T> ai = mv(st[S_AX * cap + c], st[S_AY * cap + c], st[S_AZ * cap + c]);
  const V<T> gi = mv(st[S_GX * cap + c], st[S_GY * cap + c], st[S_GZ * cap + c]);
  vv.x += hdt * (af.x - ai.x);
  vv.y += hdt * (af.y - ai.y);
  vv.z += hdt * (af.z - ai.z);
  ww.x += hdt * (gf.x - gi.x);
  ww.y += hdt * (gf.y - gi.y);
  ww.z += hdt * (gf.z - gi.z);
  T *o = st_out;
  o[S_RX * cap + c] = rr.x;
  o[S_RY * cap + c] = rr.y;
  o[S_RZ * cap + c] = rr.z;
  o[S_VX * cap + c] = vv.x;
  o[S_VY * cap + c] = vv.y;
  o[S_VZ * cap + c] = vv.z;
  o[S_WX * cap + c] = ww.x;
  o[S_WY * cap + c] = ww.y;
  o[S_WZ * cap + c] = ww.z;
  for (int k = 0; k < 9; k++) o[(S_R00 + k) * cap + c] = R[k];
  o[S_AX * cap + c] = af.x;
  o[S_AY * cap + c] = af.y;
  o[S_AZ * cap + c] = af.z;
  o[S_GX * cap + c] = gf.x;
  o[S_GY * cap + c] = gf.y;
  o[S_GZ * cap + c] = gf.z;
}

}  // namespace avf
