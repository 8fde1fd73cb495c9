// dem_window.hip -- forward, tangent and adjoint kernels of the differentiable window.
//
// Over a window the contact list is frozen: every directed sphere pair (l <- m) closer than
// sep + skin at the window start, stored per clump in the order force_part.cu accumulates
// (pairing sph_id major, then clump cell offset, then reference index).  Each step is
//   k_win_predict   init_Verlet + sphere centres -> stage (x', v*, w*, R', spheres)
//   k_win_force     boundary modules + listed contacts + get_accel + finish_Verlet
// and the adjoint of one step (from the stored stage and torque of the step) is
//   k_win_adj_force reverse of finish, accel, walls and the listed contacts of clump l;
//                   the adjoint owed to the neighbour m of each entry goes to ebuf
//   k_win_adj_gather  per clump m: its own stage adjoint + the ebuf entries that name it
//                   (transposed list, fixed order), then the reverse of the sphere centres
//                   and of init_Verlet.
// All adjoint sums are sequential in a fixed order: no float atomics.
//
// Supported: the avalanche phase (time > 0) with the hopper x / y walls, flume side walls,
// flume floor, deposit floor and particle contacts; isotropic clump inertia (clump_mode
// 1); no clump leaving the domain (x > side.x) within the window.  The host checks these.
#include "dem_window.h"

namespace avf {

constexpr int kWinMaxNeigh = 48;   // clumps within the skin of a clump at the window start
constexpr int kWinMaxList = 128;   // listed sphere pairs per clump
constexpr int kWinMaxIn = 160;     // listed entries naming a clump
constexpr int kStageW = 30;        // x'(3) v*(3) w*(3) R'(9) spheres(12)
enum WinStage { W_X = 0, W_V = 3, W_W = 6, W_R = 9, W_S = 18 };
constexpr int kEbuf = 12;          // adjoint owed to m: x'(3), v*(3), w*(3), its sphere (3)

// Window diagnostics (device).
struct WinDiag {
  unsigned long long list_overflow, in_overflow, max_list, max_in, frozen, aniso, left, pad;
};

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
__device__ __forceinline__ T wpar(const WinParams &w, int set, int k);
template <>
__device__ __forceinline__ double wpar<double>(const WinParams &w, int set, int k) {
  return w.v[set][k];
}
template <>
__device__ __forceinline__ Dual wpar<Dual>(const WinParams &w, int set, int k) {
  return mkd(w.v[set][k], w.d[set][k]);
}

// Clump cell of l by division (force_part.cu:77-85).
__device__ __forceinline__ void ref_cell_of(const InstParams &p, double x, double y, double z,
                                            int64_t &li, int64_t &lj, int64_t &lk) {
  li = (int64_t)((x - p.rr_min[0]) / p.cell_side[0]);
  if (li == p.ncell[0]) li--;
  lj = (int64_t)((y - p.rr_min[1]) / p.cell_side[1]);
  if (lj == p.ncell[1]) lj--;
  lk = (int64_t)((z - p.rr_min[2]) / p.cell_side[2]);
  if (lk == p.ncell[2]) lk--;
}

// Whether force_part visits clump m from clump l (cell li, lj, lk): m's clump cell
// (cell_locate of its centre, cells.cu:69-91) must be one of the 27 around l's
// (force_part.cu:88-96).  Clumps below or beyond the clump-cell grid are not visited.
__device__ __forceinline__ bool ref_visits(const InstParams &p, int64_t li, int64_t lj, int64_t lk,
                                           double mx, double my, double mz) {
  int64_t ai, aj, ak;
  if (!ref_cell_ijk(p, make_d3(mx, my, mz), ai, aj, ak)) return false;
  const int64_t di = ai - li, dj = aj - lj, dk = ak - lk;
  return di >= -1 && di <= 1 && dj >= -1 && dj <= 1 && dk >= -1 && dk <= 1;
}

// The next step's parameters: a copy of P_in with the clump-cell regrid (when due) and the
// step count of the plain step (k_finalize), from the k_cellreduce accumulators.
__global__ void k_win_finalize(const InstParams *__restrict__ P_in, InstParams *__restrict__ P_out,
                               long long *__restrict__ acc, int32_t ninst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ninst) return;
  InstParams p = P_in[i];
  finalize_params(p, acc + 8 * i, 0, 1);
  P_out[i] = p;
}

// ------------------------------------------------------------- list construction ----

// Frozen list of one clump from the binned current state: sphere pairs within sep + skin,
// grouped by neighbour clump (sorted by the reference key) and emitted pairing-major.
// Also records the start sphere centres (for the skin check) and flags frozen clumps and
// anisotropic inertia, which the window does not support.
__global__ __launch_bounds__(64) void k_win_list(
    const double *__restrict__ st, const double *__restrict__ sta,
    const double *__restrict__ stage, const InstParams *__restrict__ P,
    const int32_t *__restrict__ inst_of, const int32_t *__restrict__ orig,
    const uint32_t *__restrict__ starts,
    const double4 *__restrict__ spos, const uint32_t *__restrict__ sid, double skin,
    int32_t *__restrict__ list, int32_t *__restrict__ cnt_out, double *__restrict__ sph0,
    double *__restrict__ dmax2, WinDiag *__restrict__ wd, int64_t n, int64_t cap) {
  __shared__ int32_t nm[kWinMaxNeigh][64];
  __shared__ uint32_t nk[kWinMaxNeigh][64];
  __shared__ uint16_t nmask[kWinMaxNeigh][64];
  const int lane = threadIdx.x;
  const int64_t c = (int64_t)xcd_block(blockIdx.x, gridDim.x) * 64 + lane;
  if (c >= n) return;
  const InstParams &p = P[inst_of[c]];
  const double *gl = stage + c * kStageDoubles;
  const d3 rr = make_d3(gl[G_RR], gl[G_RR + 1], gl[G_RR + 2]);
  const double rad = gl[G_RAD];
  if (frozen(p.time, st[S_RX * cap + c], p.side_x) || !(p.time > 0.0))
    atomicAdd(&wd->frozen, 1ull);
  if (!(sta[C_IX * cap + c] == sta[C_IY * cap + c] && sta[C_IY * cap + c] == sta[C_IZ * cap + c]))
    atomicAdd(&wd->aniso, 1ull);
  double R0[9];
  for (int k = 0; k < 9; k++) R0[k] = gl[G_R + k];
  d3 sph[4];
  sphere_centres(rr, R0, sta[C_RAD0 * cap + c] / p.sphere_div, sph);
  for (int s = 0; s < 4; s++) {
    sph0[c * 12 + 3 * s] = sph[s].x;
    sph0[c * 12 + 3 * s + 1] = sph[s].y;
    sph0[c * 12 + 3 * s + 2] = sph[s].z;
  }
  dmax2[c] = 0.0;
  int cnt = 0;
  bool overflow = false;
  int64_t li = (int64_t)((rr.x - p.rr_min[0]) / p.cell_side[0]);
  if (li == p.ncell[0]) li--;
  int64_t lj = (int64_t)((rr.y - p.rr_min[1]) / p.cell_side[1]);
  if (lj == p.ncell[1]) lj--;
  int64_t lk = (int64_t)((rr.z - p.rr_min[2]) / p.cell_side[2]);
  if (lk == p.ncell[2]) lk--;
  const int nx = p.sg_n[0], ny = p.sg_n[1], nz = p.sg_n[2];
  // sphere cells (side >= 2 r_max) must cover sep + skin: widen the sweep by the skin
  const int reach = 1 + (int)ceil(skin * p.sg_inv_h);
  for (int s = 0; s < 4; s++) {
    const d3 q = sph[s];
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
          // pairing sph_id - 1 = 4 * (sphere of m) + (sphere of l)
          const uint32_t bit = 1u << (((id & 3u) << 2) | (uint32_t)s);
          int e = 0;
          for (; e < cnt; e++)
            if (nm[e][lane] == (int32_t)cm) break;
          if (e < cnt) {
            nmask[e][lane] = (uint16_t)(nmask[e][lane] | bit);
            continue;
          }
          const double *g = stage + cm * kStageDoubles;
          // clumps outside the reference's 27 clump cells (or out of its grid) go last
          int64_t ai, aj, ak;
          uint32_t off = 27u;
          if (ref_cell_ijk(p, make_d3(g[G_RR], g[G_RR + 1], g[G_RR + 2]), ai, aj, ak)) {
            const int64_t di = ai - li, dj = aj - lj, dk = ak - lk;
            if (di >= -1 && di <= 1 && dj >= -1 && dj <= 1 && dk >= -1 && dk <= 1)
              off = (uint32_t)((di + 1) * 9 + (dj + 1) * 3 + (dk + 1));
          }
          const uint32_t key = (off << 27) | (uint32_t)orig[cm];
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
  if (overflow) atomicAdd(&wd->list_overflow, 1ull);
  diag_max(&wd->max_list, (unsigned long long)out);
}

// Transposed list: for every clump m, the entries (l * kWinMaxList + e) that name it.
__global__ void k_win_in_fill(const int32_t *__restrict__ list, const int32_t *__restrict__ cnt,
                              int32_t *__restrict__ in, int32_t *__restrict__ icnt,
                              WinDiag *__restrict__ wd, int64_t n) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  for (int e = 0; e < cnt[l]; e++) {
    const int32_t m = list[l * kWinMaxList + e] >> 4;
    const int slot = atomicAdd(&icnt[m], 1);
    if (slot < kWinMaxIn)
      in[(int64_t)m * kWinMaxIn + slot] = (int32_t)(l * kWinMaxList + e);
    else
      atomicAdd(&wd->in_overflow, 1ull);
  }
}

// Sorts each transposed list by (reference index of l, e), so the adjoint gather sums in
// a fixed order that does not depend on the storage order of the clumps.
__global__ void k_win_in_sort(int32_t *__restrict__ in, int32_t *__restrict__ icnt,
                              const int32_t *__restrict__ orig, WinDiag *__restrict__ wd,
                              int64_t n) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  int k = icnt[m];
  diag_max(&wd->max_in, (unsigned long long)k);
  if (k > kWinMaxIn) k = icnt[m] = kWinMaxIn;
  int32_t *a = in + m * kWinMaxIn;
  auto key = [&](int32_t slot) -> int64_t {
    return (int64_t)orig[slot / kWinMaxList] * kWinMaxList + slot % kWinMaxList;
  };
  for (int i = 1; i < k; i++) {
    const int32_t v = a[i];
    const int64_t kv = key(v);
    int j = i - 1;
    while (j >= 0 && key(a[j]) > kv) {
      a[j + 1] = a[j];
      j--;
    }
    a[j + 1] = v;
  }
}

// Touching sphere pairs of a binned state (stage, spos, sid, starts of the window's own
// binning) that the frozen list lacks: one lane per sphere of clump l, 27 sphere cells.
__global__ void k_win_missed(const double *__restrict__ stage, const InstParams *__restrict__ P,
                             const int32_t *__restrict__ inst_of, const uint32_t *__restrict__ starts,
                             const double4 *__restrict__ spos, const uint32_t *__restrict__ sid,
                             const int32_t *__restrict__ list, const int32_t *__restrict__ cnt,
                             unsigned long long *__restrict__ missed, int64_t nsph) {
  const int64_t i = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= nsph) return;
  const int64_t c = i >> 2;
  const int s = (int)(i & 3);
  const InstParams &p = P[inst_of[c]];
  const double *gl = stage + c * kStageDoubles;
  const double rad = gl[G_RAD];
  const d3 q = rec_sphere(gl, rad / p.sphere_div, s);
  const int nx = p.sg_n[0], ny = p.sg_n[1], nz = p.sg_n[2];
  const int ix = grid_index(q.x, p.sg_lo[0], p.sg_inv_h, nx);
  const int iy = grid_index(q.y, p.sg_lo[1], p.sg_inv_h, ny);
  const int iz = grid_index(q.z, p.sg_lo[2], p.sg_inv_h, nz);
  const int x0 = ix > 0 ? ix - 1 : 0, x1 = ix + 1 < nx ? ix + 1 : nx - 1;
  const int k = cnt[c];
  unsigned long long miss = 0;
  for (int z = iz - 1; z <= iz + 1; z++) {
    if (z < 0 || z >= nz) continue;
    for (int y = iy - 1; y <= iy + 1; y++) {
      if (y < 0 || y >= ny) continue;
      const int64_t row = p.sg_base + (int64_t)nx * (y + (int64_t)ny * z);
      for (uint32_t j = starts[row + x0]; j < starts[row + x1 + 1]; j++) {
        const uint32_t id = sid[j];
        const int64_t cm = (int64_t)(id >> 2);
        if (cm == c) continue;
        const double4 o = spos[j];
        const double sep = rad + o.w;
        const double dx = o.x - q.x, dy = o.y - q.y, dz = o.z - q.z;
        if (!(dx * dx + dy * dy + dz * dz < sep * sep)) continue;
        const int32_t ent = (int32_t)((cm << 4) | (int64_t)(((id & 3u) << 2) | (uint32_t)s));
        bool found = false;
        for (int e = 0; e < k && !found; e++) found = list[c * kWinMaxList + e] == ent;
        if (!found) miss++;
      }
    }
  }
  if (miss) atomicAdd(missed, miss);
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

// State (field-major SoA of T) -> stage (AoS of T, kStageW per clump).
template <typename T>
__global__ __launch_bounds__(256) void k_win_predict(const T *__restrict__ st, const double *__restrict__ sta,
                              const InstParams *__restrict__ P, const int32_t *__restrict__ inst_of,
                              T *__restrict__ stage, int64_t n, int64_t cap) {
  const int64_t c = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const InstParams &p = P[inst_of[c]];
  T S[S_NFIELDS];
  for (int f = 0; f < S_NFIELDS; f++) S[f] = st[f * cap + c];
  win_predict_one(S, p.dt, p.hdt, sta[C_RAD0 * cap + c] / p.sphere_div, stage + c * kStageW);
}

// Wall contact of one sphere given its contact point, normal and compression.
template <typename T>
__device__ __forceinline__ void win_wall(V<T> rr, V<T> vv, V<T> ww, V<T> rrc, V<T> n, T comp,
                                         T km, T gm, T mu, T cns, double dt, double mass,
                                         V<T> &ff, V<T> &tq) {
  V<T> vc = pvel(vv, ww, rrc, rr);
  V<T> f;
  if (contact_force(mv(-vc.x, -vc.y, -vc.z), n, comp, km, gm, mu, cns, dt, mass, false, f))
    add_ft(ff, tq, rrc, rr, f);
}

// Hopper x / y walls (force_hop.cu:9-244, nw = 0, 1; the gate is open) and flume side walls
// (force_walls.cu:9-199), in the reference order.
template <typename T>
__device__ __forceinline__ void win_walls_pre(const InstParams &p, int mods, V<T> rr, V<T> vv,
                                              V<T> ww, const T *sph, double ri, double mass,
                                              const WinParams &wp, V<T> &ff, V<T> &tq) {
  const T kap = wpar<T>(wp, WS_PW, WP_KAPPA), gam = wpar<T>(wp, WS_PW, WP_GAMMA);
  const T mu = wpar<T>(wp, WS_PW, WP_MU), cns = wpar<T>(wp, WS_PW, WP_FR);
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
    // force_walls.cu: without inter_flag the side walls are frictionless and undamped
    // once the gate is open
    const T zero = lift<T>(0.0);
    const T gmw = p.inter_flag ? gm : zero, muw = p.inter_flag ? mu : zero;
    const T cnsw = p.inter_flag ? cns : zero;
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
      if (touch) win_wall(rr, vv, ww, rrc, nv, comp, km, gmw, muw, cnsw, p.dt, mass, ff, tq);
    }
  }
}

// Flume floor (force_flume.cu:10-182) and deposit floor (force_dep.cu:10-178).
template <typename T>
__device__ __forceinline__ void win_walls_post(const InstParams &p, int mods, V<T> rr, V<T> vv,
                                               V<T> ww, const T *sph, double ri, double mass,
                                               const WinParams &wp, V<T> &ff, V<T> &tq) {
  const T kap = wpar<T>(wp, WS_PW, WP_KAPPA), gam = wpar<T>(wp, WS_PW, WP_GAMMA);
  const T mu = wpar<T>(wp, WS_PW, WP_MU), cns = wpar<T>(wp, WS_PW, WP_FR);
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

// Listed contacts, get_accel and finish_Verlet for one clump: stage -> new state in st_out
// (a buffer distinct from st).  Optional outputs: tq_out keeps the total torque of the step
// (for the adjoint of get_accel); sph0 / dmax2 track the largest sphere displacement from
// the window start (skin check); wd counts clumps that reach x > side.x.
template <typename T>
__global__ __launch_bounds__(64) void k_win_force(
    const T *__restrict__ st, T *__restrict__ st_out, const double *__restrict__ sta,
    const T *__restrict__ stage, const InstParams *__restrict__ P,
    const WinParams *__restrict__ WP, const int32_t *__restrict__ inst_of,
    const int32_t *__restrict__ list, const int32_t *__restrict__ cnt, T *__restrict__ tq_out,
    const double *__restrict__ sph0, double *__restrict__ dmax2, WinDiag *__restrict__ wd,
    int64_t n, int64_t cap) {
  const int64_t c = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int inst = inst_of[c];
  const InstParams &p = P[inst];
  const WinParams &wp = WP[inst];
  const T *g = stage + c * kStageW;
  const V<T> rr = ld3(g + W_X);
  V<T> vv = ld3(g + W_V), ww = ld3(g + W_W);
  const double rad = sta[C_RAD * cap + c], rho = sta[C_RHO * cap + c];
  const double massw = kSphFact * rad * rad * rad * rho;  // force_hop.cu:45 order
  const int mods = p.modules_aval;
  V<T> ff = mv(lift<T>(0.0), lift<T>(0.0), lift<T>(0.0)), tq = ff;
  win_walls_pre(p, mods, rr, vv, ww, g + W_S, rad, massw, wp, ff, tq);
  if (mods & AVF_MOD_PP) {
    const double massl = kSphFact * rho * rad * rad * rad;  // force_part.cu order
    const T kap = wpar<T>(wp, WS_PP, WP_KAPPA), gam = wpar<T>(wp, WS_PP, WP_GAMMA);
    const T mu = wpar<T>(wp, WS_PP, WP_MU), cns = wpar<T>(wp, WS_PP, WP_FR);
    int64_t li, lj, lk;
    ref_cell_of(p, val(rr.x), val(rr.y), val(rr.z), li, lj, lk);
    const int k = cnt[c];
    for (int e = 0; e < k; e++) {
      const int32_t ent = list[c * kWinMaxList + e];
      const int64_t m = ent >> 4;
      const int pr = ent & 15, sdiv = pr >> 2, smod = pr & 3;
      const T *gm_ = stage + m * kStageW;
      if (!ref_visits(p, li, lj, lk, val(gm_[W_X]), val(gm_[W_X + 1]), val(gm_[W_X + 2]))) continue;
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
  if (tq_out) {
    tq_out[3 * c + 0] = tq.x;
    tq_out[3 * c + 1] = tq.y;
    tq_out[3 * c + 2] = tq.z;
  }
  if (dmax2) {
    double d2 = dmax2[c];
    for (int s = 0; s < 4; s++) {
      const double dx = val(g[W_S + 3 * s]) - sph0[c * 12 + 3 * s];
      const double dy = val(g[W_S + 3 * s + 1]) - sph0[c * 12 + 3 * s + 1];
      const double dz = val(g[W_S + 3 * s + 2]) - sph0[c * 12 + 3 * s + 2];
      const double q = dx * dx + dy * dy + dz * dz;
      d2 = q > d2 ? q : d2;
    }
    dmax2[c] = d2;
  }
  if (wd && frozen(p.time, val(rr.x), p.side_x)) atomicAdd(&wd->left, 1ull);
  // get_accel.cu:37-71 (isotropic inertia: the gyroscopic terms vanish)
  const double mass = sta[C_MASS * cap + c], I = sta[C_IX * cap + c];
  T R[9];
  for (int k = 0; k < 9; k++) R[k] = g[W_R + k];
  double cd = p.drag * rad * rad / mass;
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

// ------------------------------------------------------------------- adjoint step ----

__device__ __forceinline__ V<double> z3() { return mv(0.0, 0.0, 0.0); }
__device__ __forceinline__ void acc3(V<double> &a, V<double> b) {
  a.x += b.x;
  a.y += b.y;
  a.z += b.z;
}

// Reverse of one wall contact given the clump's force and torque adjoints (Fb, Tb).
// Accumulates the adjoints of x', v*, w*, the contact point, the compression and the law
// parameters; returns whether the contact was repulsive.
__device__ __forceinline__ bool wall_bwd(V<double> rr, V<double> vv, V<double> ww, V<double> rrc,
                                         V<double> n, double comp, double km, double gm,
                                         double mu, double cns, double dt, double mass,
                                         V<double> Fb, V<double> Tb, V<double> &xb,
                                         V<double> &vb, V<double> &wb, V<double> &rrcb,
                                         double &compb, double &kmb, double &gmb, double &mub) {
  V<double> vc = pvel(vv, ww, rrc, rr), f;
  V<double> dv = mv(-vc.x, -vc.y, -vc.z);
  if (!contact_force(dv, n, comp, km, gm, mu, cns, dt, mass, false, f)) return false;
  V<double> fb = add_ft_bwd(rrc, rr, f, Fb, Tb, rrcb, xb);
  V<double> dvb = z3(), nb = z3();
  contact_force_bwd(dv, n, comp, km, gm, mu, cns, dt, mass, false, fb, dvb, nb, compb, kmb, gmb,
                    mub);
  pvel_bwd(ww, rrc, rr, mv(-dvb.x, -dvb.y, -dvb.z), vb, wb, rrcb, xb);
  return true;
}

// Reverse of finish_Verlet, get_accel, the boundary modules and the listed contacts of
// clump l.  Sb: adjoint of the post-step state (SoA, StateField order).  Writes the stage
// adjoint of l (sbar), the adjoint owed to each listed neighbour (ebuf, with eflag set for
// repulsive entries), the finish_Verlet part of the pre-step a_i / alpha_i adjoints
// (Sb_prev) and accumulates the contact-law parameter adjoints (tbar, per clump).
__global__ __launch_bounds__(64) void k_win_adj_force(
    const double *__restrict__ Sb, double *__restrict__ Sb_prev, const double *__restrict__ sta,
    const double *__restrict__ stage, const double *__restrict__ tqs,
    const InstParams *__restrict__ P, const WinParams *__restrict__ WP,
    const int32_t *__restrict__ inst_of, const int32_t *__restrict__ list,
    const int32_t *__restrict__ cnt, double *__restrict__ sbar, double *__restrict__ ebuf,
    uint8_t *__restrict__ eflag, double *__restrict__ tbar, int64_t n, int64_t cap) {
  const int64_t c = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int inst = inst_of[c];
  const InstParams &p = P[inst];
  const WinParams &wp = WP[inst];
  const double *g = stage + c * kStageW;
  const V<double> rr = ld3(g + W_X), vv = ld3(g + W_V), ww = ld3(g + W_W);
  double R[9];
  for (int k = 0; k < 9; k++) R[k] = g[W_R + k];
  const double rad = sta[C_RAD * cap + c], rho = sta[C_RHO * cap + c];
  const double mass = sta[C_MASS * cap + c], I = sta[C_IX * cap + c];
  const double hdt = p.hdt, dt = p.dt;
  const V<double> vb_new = mv(Sb[S_VX * cap + c], Sb[S_VY * cap + c], Sb[S_VZ * cap + c]);
  const V<double> wb_new = mv(Sb[S_WX * cap + c], Sb[S_WY * cap + c], Sb[S_WZ * cap + c]);
  const V<double> ab_new = mv(Sb[S_AX * cap + c], Sb[S_AY * cap + c], Sb[S_AZ * cap + c]);
  const V<double> gb_new = mv(Sb[S_GX * cap + c], Sb[S_GY * cap + c], Sb[S_GZ * cap + c]);
  double Rb[9];
  for (int k = 0; k < 9; k++) Rb[k] = Sb[(S_R00 + k) * cap + c];
  V<double> xb = mv(Sb[S_RX * cap + c], Sb[S_RY * cap + c], Sb[S_RZ * cap + c]);
  V<double> vb = vb_new, wb = wb_new;
  // finish_Verlet: v'' = v* + hdt (a' - a_i), w'' = w* + hdt (alpha' - alpha_i); a', alpha'
  // are also stored as the next a_i, alpha_i
  const V<double> afb = mv(ab_new.x + hdt * vb_new.x, ab_new.y + hdt * vb_new.y,
                           ab_new.z + hdt * vb_new.z);
  const V<double> gfb = mv(gb_new.x + hdt * wb_new.x, gb_new.y + hdt * wb_new.y,
                           gb_new.z + hdt * wb_new.z);
  Sb_prev[S_AX * cap + c] = -hdt * vb_new.x;
  Sb_prev[S_AY * cap + c] = -hdt * vb_new.y;
  Sb_prev[S_AZ * cap + c] = -hdt * vb_new.z;
  Sb_prev[S_GX * cap + c] = -hdt * wb_new.x;
  Sb_prev[S_GY * cap + c] = -hdt * wb_new.y;
  Sb_prev[S_GZ * cap + c] = -hdt * wb_new.z;
  // get_accel: a' = F / M - cd |v*| v* - g e_z ;  alpha' = R' (R'^T T / I)
  const V<double> Fb = mv(afb.x / mass, afb.y / mass, afb.z / mass);
  const double cd = p.drag * rad * rad / mass;
  const double vmag = sqrt(vv.x * vv.x + vv.y * vv.y + vv.z * vv.z);
  if (vmag > 0.0) {
    const double va = vv.x * afb.x + vv.y * afb.y + vv.z * afb.z;
    vb.x += -cd * (vmag * afb.x + vv.x * va / vmag);
    vb.y += -cd * (vmag * afb.y + vv.y * va / vmag);
    vb.z += -cd * (vmag * afb.z + vv.z * va / vmag);
  }
  const double Tq[3] = {tqs[3 * c], tqs[3 * c + 1], tqs[3 * c + 2]};
  // B = R^T T / I ; alpha' = R B
  double Bv[3];
  for (int j = 0; j < 3; j++) Bv[j] = (Tq[0] * R[j] + Tq[1] * R[3 + j] + Tq[2] * R[6 + j]) / I;
  const double gfa[3] = {gfb.x, gfb.y, gfb.z};
  double Bb[3];
  for (int j = 0; j < 3; j++) {
    double bb = 0.0;
    for (int i = 0; i < 3; i++) {
      bb += R[3 * i + j] * gfa[i];
      Rb[3 * i + j] += gfa[i] * Bv[j];
    }
    Bb[j] = bb / I;
  }
  double Tba[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      Tba[i] += R[3 * i + j] * Bb[j];
      Rb[3 * i + j] += Tq[i] * Bb[j];
    }
  const V<double> Tb = mv(Tba[0], Tba[1], Tba[2]);
  // boundary modules
  const int mods = p.modules_aval;
  const double massw = kSphFact * rad * rad * rad * rho;
  const double kap_w = wp.v[WS_PW][WP_KAPPA], gam_w = wp.v[WS_PW][WP_GAMMA];
  const double mu_w = wp.v[WS_PW][WP_MU], fr_w = wp.v[WS_PW][WP_FR];
  const double km_w = 0.5 * kap_w * massw, gm_w = 0.5 * gam_w * massw;
  double kmb_w = 0.0, gmb_w = 0.0, mub_w = 0.0, kmb_p = 0.0, gmb_p = 0.0, mub_p = 0.0;
  V<double> sb[4] = {z3(), z3(), z3(), z3()};
  for (int ns = 0; ns < 4; ns++) {
    const V<double> r = ld3(g + W_S + 3 * ns);
    V<double> sl = z3();  // adjoint of this sphere centre
    if (mods & AVF_MOD_HOP) {
      if (!(r.z < p.hop_height || r.x > p.hop_width)) {
        for (int nw = 0; nw < 2; nw++) {
          const double lo = nw ? p.yhop_l : 0.0, hi = nw ? p.yhop_r : p.hop_width;
          const double coord = nw ? r.y : r.x;
          int touch = 0;
          double comp = 0.0;
          if (coord - lo <= rad) {
            touch = 1;
            comp = rad - (coord - lo);
          }
          if (hi - coord <= rad) {
            touch = 2;
            comp = rad - (hi - coord);
          }
          if (!touch) continue;
          const double wallc = touch == 1 ? lo : hi;
          const V<double> rrc = nw ? mv(r.x, wallc, r.z) : mv(wallc, r.y, r.z);
          const double sgn = touch == 1 ? -1.0 : 1.0;
          const V<double> nv = nw ? mv(0.0, sgn, 0.0) : mv(sgn, 0.0, 0.0);
          V<double> rrcb = z3();
          double compb = 0.0;
          if (wall_bwd(rr, vv, ww, rrc, nv, comp, km_w, gm_w, mu_w, fr_w, dt, massw, Fb, Tb, xb,
                       vb, wb, rrcb, compb, kmb_w, gmb_w, mub_w)) {
            // comp = rad - (coord - lo) or rad - (hi - coord); rrc copies the other two
            const double dcomp = touch == 1 ? -compb : compb;
            if (nw) {
              sl.x += rrcb.x;
              sl.y += dcomp;
              sl.z += rrcb.z;
            } else {
              sl.x += dcomp;
              sl.y += rrcb.y;
              sl.z += rrcb.z;
            }
          }
        }
      }
    }
    if (mods & AVF_MOD_WALLS) {
      const double qx = r.x * p.cost - r.z * p.sint;
      if (qx > p.lfunnel) {
        const double yl = p.dy2, yr = p.side_y + p.dy2;
        int touch = 0;
        double comp = 0.0;
        if (r.y - yl <= rad) {
          touch = 1;
          comp = rad - (r.y - yl);
        }
        if (yr - r.y <= rad) {
          touch = 2;
          comp = rad - (yr - r.y);
        }
        if (touch) {
          const V<double> rrc = mv(r.x, touch == 1 ? yl : yr, r.z);
          const V<double> nv = mv(0.0, touch == 1 ? -1.0 : 1.0, 0.0);
          V<double> rrcb = z3();
          double compb = 0.0, gdrop = 0.0, mdrop = 0.0;
          const bool il = p.inter_flag != 0;
          if (wall_bwd(rr, vv, ww, rrc, nv, comp, km_w, il ? gm_w : 0.0, il ? mu_w : 0.0,
                       il ? fr_w : 0.0, dt, massw, Fb, Tb, xb, vb, wb, rrcb, compb, kmb_w,
                       il ? gmb_w : gdrop, il ? mub_w : mdrop)) {
            sl.x += rrcb.x;
            sl.y += touch == 1 ? -compb : compb;
            sl.z += rrcb.z;
          }
        }
      }
    }
    if (mods & AVF_MOD_FLUME) {
      const double sint = p.sint, cost = p.cost, tant = sint / cost;
      const double rect = -tant * r.x - r.z, dist = ::fabs(cost * rect);
      if (dist <= rad) {
        const V<double> nv = mv(-sint, 0.0, -cost);
        const V<double> rrc = mv(r.x - dist * sint, r.y, r.z - dist * cost);
        V<double> rrcb = z3();
        double compb = 0.0;
        if (wall_bwd(rr, vv, ww, rrc, nv, rad - dist, km_w, gm_w, mu_w, fr_w, dt, massw, Fb, Tb,
                     xb, vb, wb, rrcb, compb, kmb_w, gmb_w, mub_w)) {
          acc3(sl, rrcb);
          const double distb = -compb - rrcb.x * sint - rrcb.z * cost;
          const double rectb = distb * (cost * rect < 0.0 ? -cost : cost);
          sl.x += -tant * rectb;
          sl.z += -rectb;
        }
      }
    }
    if (mods & AVF_MOD_DEP) {
      const double sind = p.dep_sin, cosd = p.dep_cos, tand = sind / cosd;
      const double rect = tand * (r.x - p.dep_x0 - p.dep_x1) - (r.z - p.dep_z0);
      const double dist = cosd * (-rect);
      if (dist <= rad) {
        const double rx = r.x + dist * sind;
        const V<double> nv = mv(sind, 0.0, -cosd);
        const V<double> rrc = mv(rx, r.y, tand * (rx - p.dep_x0 - p.dep_x1) + p.dep_z0);
        V<double> rrcb = z3();
        double compb = 0.0;
        if (wall_bwd(rr, vv, ww, rrc, nv, rad - dist, km_w, gm_w, mu_w, fr_w, dt, massw, Fb, Tb,
                     xb, vb, wb, rrcb, compb, kmb_w, gmb_w, mub_w)) {
          const double rxb = rrcb.x + tand * rrcb.z;
          sl.x += rxb;
          sl.y += rrcb.y;
          const double distb = rxb * sind - compb;
          const double rectb = -cosd * distb;
          sl.x += tand * rectb;
          sl.z += -rectb;
        }
      }
    }
    for (int s = 0; s < 4; s++)
      if (s == ns) acc3(sb[s], sl);
  }
  // listed contacts (l <- m)
  if (mods & AVF_MOD_PP) {
    const double massl = kSphFact * rho * rad * rad * rad;
    const double kap = wp.v[WS_PP][WP_KAPPA], gam = wp.v[WS_PP][WP_GAMMA];
    const double mu = wp.v[WS_PP][WP_MU], fr = wp.v[WS_PP][WP_FR];
    int64_t li, lj, lk;
    ref_cell_of(p, rr.x, rr.y, rr.z, li, lj, lk);
    const int k = cnt[c];
    for (int e = 0; e < k; e++) {
      const int64_t slot = c * kWinMaxList + e;
      const int32_t ent = list[slot];
      const int64_t m = ent >> 4;
      const int pr = ent & 15, sdiv = pr >> 2, smod = pr & 3;
      const double *gm_ = stage + m * kStageW;
      const double rad2 = sta[C_RAD * cap + m], rho2 = sta[C_RHO * cap + m];
      const V<double> r1 = ld3(g + W_S + 3 * smod), r2 = ld3(gm_ + W_S + 3 * sdiv);
      const double sep = rad + rad2;
      const V<double> drr = mv(r2.x - r1.x, r2.y - r1.y, r2.z - r1.z);
      const double dist2 = drr.x * drr.x + drr.y * drr.y + drr.z * drr.z;
      uint8_t active = 0;
      if (dist2 < sep * sep && ref_visits(p, li, lj, lk, gm_[W_X], gm_[W_X + 1], gm_[W_X + 2])) {
        const double massm = kSphFact * rho2 * rad2 * rad2 * rad2;
        const double meff = 0.5 * (massm + massl);
        const double km = 0.5 * kap * meff, gmm = 0.5 * gam * meff;
        const double dist = sqrt(dist2), comp = sep - dist;
        const V<double> nv = mv(drr.x / dist, drr.y / dist, drr.z / dist);
        const V<double> rrc = mv(r1.x + rad * nv.x, r1.y + rad * nv.y, r1.z + rad * nv.z);
        const V<double> xm = ld3(gm_ + W_X), vm = ld3(gm_ + W_V), wm = ld3(gm_ + W_W);
        const V<double> vl = pvel(vv, ww, rrc, rr), vmc = pvel(vm, wm, rrc, xm);
        const V<double> dvvc = mv(vmc.x - vl.x, vmc.y - vl.y, vmc.z - vl.z);
        V<double> f;
        if (contact_force(dvvc, nv, comp, km, gmm, mu, fr, dt, massl, true, f)) {
          active = 1;
          V<double> rrcb = z3(), dvb = z3(), nb = z3();
          double compb = 0.0, kmb = 0.0, gmb = 0.0, mub = 0.0;
          const V<double> fb = add_ft_bwd(rrc, rr, f, Fb, Tb, rrcb, xb);
          contact_force_bwd(dvvc, nv, comp, km, gmm, mu, fr, dt, massl, true, fb, dvb, nb, compb,
                            kmb, gmb, mub);
          kmb_p += kmb * 0.5 * meff;
          gmb_p += gmb * 0.5 * meff;
          mub_p += mub;
          // dvvc = vmc - vl
          V<double> xmb = z3(), vmb = z3(), wmb = z3();
          pvel_bwd(ww, rrc, rr, mv(-dvb.x, -dvb.y, -dvb.z), vb, wb, rrcb, xb);
          pvel_bwd(wm, rrc, xm, dvb, vmb, wmb, rrcb, xmb);
          // rrc = r1 + rad n
          V<double> r1b = rrcb;
          nb.x += rad * rrcb.x;
          nb.y += rad * rrcb.y;
          nb.z += rad * rrcb.z;
          // n = drr / dist, comp = sep - dist, dist = |drr|
          const double nd = nb.x * drr.x + nb.y * drr.y + nb.z * drr.z;
          const double distb = -nd / (dist * dist) - compb;
          const V<double> drrb = mv(nb.x / dist + distb * drr.x / dist,
                                    nb.y / dist + distb * drr.y / dist,
                                    nb.z / dist + distb * drr.z / dist);
          // drr = r2 - r1
          r1b.x -= drrb.x;
          r1b.y -= drrb.y;
          r1b.z -= drrb.z;
          for (int s = 0; s < 4; s++)
            if (s == smod) acc3(sb[s], r1b);
          double *eb = ebuf + slot * kEbuf;
          st3(eb + 0, xmb);
          st3(eb + 3, vmb);
          st3(eb + 6, wmb);
          st3(eb + 9, drrb);
        }
      }
      eflag[slot] = active;
    }
  }
  // km = 0.5 kappa m, gm = 0.5 gamma m
  double *t = tbar + c * TB_N;
  t[TB_KPP] += kmb_p;
  t[TB_GPP] += gmb_p;
  t[TB_MPP] += mub_p;
  t[TB_KPW] += kmb_w * 0.5 * massw;
  t[TB_GPW] += gmb_w * 0.5 * massw;
  t[TB_MPW] += mub_w;
  double *o = sbar + c * kStageW;
  st3(o + W_X, xb);
  st3(o + W_V, vb);
  st3(o + W_W, wb);
  for (int k = 0; k < 9; k++) o[W_R + k] = Rb[k];
  for (int s = 0; s < 4; s++) st3(o + W_S + 3 * s, sb[s]);
}

// Per clump m: its stage adjoint plus the adjoints its listed neighbours owe it (transposed
// list, ascending), then the reverse of the sphere centres and of init_Verlet (Rodrigues
// rotation and column renormalisation included).  S: pre-step state.  Sb_prev already
// holds the finish_Verlet part of the a_i / alpha_i adjoints.
__global__ __launch_bounds__(64) void k_win_adj_gather(const double *__restrict__ S, double *__restrict__ Sb_prev,
                                 const double *__restrict__ sta, const double *__restrict__ sbar,
                                 const double *__restrict__ ebuf, const uint8_t *__restrict__ eflag,
                                 const int32_t *__restrict__ list, const int32_t *__restrict__ in,
                                 const int32_t *__restrict__ icnt, const InstParams *__restrict__ P,
                                 const int32_t *__restrict__ inst_of, int64_t n, int64_t cap) {
  const int64_t c = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const InstParams &p = P[inst_of[c]];
  const double dt = p.dt, hdt = p.hdt;
  const double *o = sbar + c * kStageW;
  V<double> xb = ld3(o + W_X), vb = ld3(o + W_V), wb = ld3(o + W_W);
  double Rb[9];
  for (int k = 0; k < 9; k++) Rb[k] = o[W_R + k];
  V<double> sb[4];
  for (int s = 0; s < 4; s++) sb[s] = ld3(o + W_S + 3 * s);
  const int k = icnt[c];
  for (int i = 0; i < k; i++) {
    const int32_t slot = in[c * kWinMaxIn + i];
    if (!eflag[slot]) continue;
    const int sdiv = (list[slot] & 15) >> 2;
    const double *eb = ebuf + (int64_t)slot * kEbuf;
    acc3(xb, ld3(eb + 0));
    acc3(vb, ld3(eb + 3));
    acc3(wb, ld3(eb + 6));
    const V<double> r2b = ld3(eb + 9);
    for (int s = 0; s < 4; s++)
      if (s == sdiv) acc3(sb[s], r2b);
  }
  // sphere centres: s_k = x' + pos R' e_k
  const double pos = sta[C_RAD0 * cap + c] / p.sphere_div;
  const double sg[4][3] = {{1, 1, 1}, {1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}};
  for (int s = 0; s < 4; s++) {
    acc3(xb, sb[s]);
    const double sbv[3] = {sb[s].x, sb[s].y, sb[s].z};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) Rb[3 * i + j] += pos * sbv[i] * sg[s][j];
  }
  // init_Verlet reverse: recompute dOmega, the rotation D and T = D R
  const V<double> ww = mv(S[S_WX * cap + c], S[S_WY * cap + c], S[S_WZ * cap + c]);
  const V<double> ang = mv(S[S_GX * cap + c], S[S_GY * cap + c], S[S_GZ * cap + c]);
  double R[9];
  for (int q = 0; q < 9; q++) R[q] = S[(S_R00 + q) * cap + c];
  const V<double> dO = mv(dt * (ww.x + hdt * ang.x), dt * (ww.y + hdt * ang.y),
                          dt * (ww.z + hdt * ang.z));
  const double mag = sqrt(dO.x * dO.x + dO.y * dO.y + dO.z * dO.z);
  double D[9];
  double nn[3] = {0.0, 0.0, 0.0}, cdO = 1.0, sdO = 0.0, om = 0.0;
  if (mag > 0.0) {
    nn[0] = dO.x / mag;
    nn[1] = dO.y / mag;
    nn[2] = dO.z / mag;
    cdO = cos(mag);
    sdO = sin(mag);
    om = 1.0 - cdO;
    D[0] = om * nn[0] * nn[0] + cdO;
    D[1] = om * nn[0] * nn[1] - sdO * nn[2];
    D[2] = om * nn[0] * nn[2] + sdO * nn[1];
    D[3] = om * nn[1] * nn[0] + sdO * nn[2];
    D[4] = om * nn[1] * nn[1] + cdO;
    D[5] = om * nn[1] * nn[2] - sdO * nn[0];
    D[6] = om * nn[2] * nn[0] - sdO * nn[1];
    D[7] = om * nn[2] * nn[1] + sdO * nn[0];
    D[8] = om * nn[2] * nn[2] + cdO;
  } else {
    D[0] = 1.0;
    D[1] = -dO.z;
    D[2] = dO.y;
    D[3] = dO.z;
    D[4] = 1.0;
    D[5] = -dO.x;
    D[6] = -dO.y;
    D[7] = dO.x;
    D[8] = 1.0;
  }
  double Tm[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      Tm[3 * i + j] = D[3 * i] * R[j] + D[3 * i + 1] * R[3 + j] + D[3 * i + 2] * R[6 + j];
  // column normalisation R'_ij = T_ij / nu_j
  double Tb[9];
  for (int j = 0; j < 3; j++) {
    const double nu = sqrt(Tm[j] * Tm[j] + Tm[3 + j] * Tm[3 + j] + Tm[6 + j] * Tm[6 + j]);
    const double r0 = Tm[j] / nu, r1 = Tm[3 + j] / nu, r2 = Tm[6 + j] / nu;
    const double dotr = r0 * Rb[j] + r1 * Rb[3 + j] + r2 * Rb[6 + j];
    Tb[j] = (Rb[j] - r0 * dotr) / nu;
    Tb[3 + j] = (Rb[3 + j] - r1 * dotr) / nu;
    Tb[6 + j] = (Rb[6 + j] - r2 * dotr) / nu;
  }
  // T = D R:  D_b = T_b R^T,  R_b = D^T T_b
  double Db[9], Rb0[9];
  for (int i = 0; i < 3; i++)
    for (int q = 0; q < 3; q++)
      Db[3 * i + q] = Tb[3 * i] * R[3 * q] + Tb[3 * i + 1] * R[3 * q + 1] + Tb[3 * i + 2] * R[3 * q + 2];
  for (int q = 0; q < 3; q++)
    for (int j = 0; j < 3; j++)
      Rb0[3 * q + j] = D[q] * Tb[j] + D[3 + q] * Tb[3 + j] + D[6 + q] * Tb[6 + j];
  // dOmega adjoint
  V<double> dOb;
  if (mag > 0.0) {
    // D = om n n^T + c I + s [n]x
    double omb = 0.0, cb = Db[0] + Db[4] + Db[8];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) omb += Db[3 * i + j] * nn[i] * nn[j];
    const double sb_ = -Db[1] * nn[2] + Db[2] * nn[1] + Db[3] * nn[2] - Db[5] * nn[0] -
                       Db[6] * nn[1] + Db[7] * nn[0];
    double nb[3];
    for (int i = 0; i < 3; i++) {
      double a = 0.0;
      for (int j = 0; j < 3; j++) a += (Db[3 * i + j] + Db[3 * j + i]) * nn[j];
      nb[i] = om * a;
    }
    nb[0] += sdO * (Db[7] - Db[5]);
    nb[1] += sdO * (Db[2] - Db[6]);
    nb[2] += sdO * (Db[3] - Db[1]);
    // om = 1 - c
    cb += -omb;
    const double magb = -sdO * cb + cdO * sb_;
    const double nd = nb[0] * nn[0] + nb[1] * nn[1] + nb[2] * nn[2];
    dOb = mv((nb[0] - nd * nn[0]) / mag + magb * nn[0], (nb[1] - nd * nn[1]) / mag + magb * nn[1],
             (nb[2] - nd * nn[2]) / mag + magb * nn[2]);
  } else {
    dOb = mv(Db[7] - Db[5], Db[2] - Db[6], Db[3] - Db[1]);
  }
  // x' = x + dt (v + hdt a), v* = v + dt a, w* = w + dt alpha, dO = dt (w + hdt alpha)
  double *so = Sb_prev;
  so[S_RX * cap + c] = xb.x;
  so[S_RY * cap + c] = xb.y;
  so[S_RZ * cap + c] = xb.z;
  so[S_VX * cap + c] = dt * xb.x + vb.x;
  so[S_VY * cap + c] = dt * xb.y + vb.y;
  so[S_VZ * cap + c] = dt * xb.z + vb.z;
  so[S_WX * cap + c] = wb.x + dt * dOb.x;
  so[S_WY * cap + c] = wb.y + dt * dOb.y;
  so[S_WZ * cap + c] = wb.z + dt * dOb.z;
  for (int q = 0; q < 9; q++) so[(S_R00 + q) * cap + c] = Rb0[q];
  so[S_AX * cap + c] += dt * hdt * xb.x + dt * vb.x;
  so[S_AY * cap + c] += dt * hdt * xb.y + dt * vb.y;
  so[S_AZ * cap + c] += dt * hdt * xb.z + dt * vb.z;
  so[S_GX * cap + c] += dt * wb.x + dt * hdt * dOb.x;
  so[S_GY * cap + c] += dt * wb.y + dt * hdt * dOb.y;
  so[S_GZ * cap + c] += dt * wb.z + dt * hdt * dOb.z;
}

// ------------------------------------------------------------------------ loss, I/O ----

// Adjoint seed: dL/d(final state) for L = w_r runout_mean + sum_b wbins_b h_b, with
// inst_w[i] = w_r / (live mass of instance i) and the caller's wbins = dL/dh (the profile
// h_b as in k_observe: Gaussian kernel of the slope-frame coordinate x' = x cos - z sin).
__global__ void k_win_loss_grad(const double *__restrict__ st, const double *__restrict__ sta,
                                const InstParams *__restrict__ P, const int32_t *__restrict__ inst_of,
                                const double *__restrict__ inst_w, const double *__restrict__ wbins,
                                int nbins, double x0, double bw, double sigma, double width,
                                double *__restrict__ Sb, int64_t n, int64_t cap) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int inst = inst_of[c];
  const InstParams &p = P[inst];
  for (int f = 0; f < S_NFIELDS; f++) Sb[f * cap + c] = 0.0;
  const double x = st[S_RX * cap + c], z = st[S_RZ * cap + c];
  if (frozen(p.time, x, p.side_x)) return;
  const double m = sta[C_MASS * cap + c], vol = m / sta[C_RHO * cap + c];
  const double xs = x * p.cost - z * p.sint;
  double g = inst_w[inst] * m;
  const double *wb = wbins + (int64_t)inst * nbins;
  for (int b = 0; b < nbins; b++) {
    const double u = xs - (x0 + (b + 0.5) * bw);
    if (fabs(u) < 8.0 * sigma) g += wb[b] * vol * gauss(u, sigma) / width * (-u / (sigma * sigma));
  }
  Sb[S_RX * cap + c] = g * p.cost;
  Sb[S_RZ * cap + c] = -g * p.sint;
}

// Per-block partial sums of per-clump vectors (width w) over reference indices
// [b * kObsBlock, (b + 1) * kObsBlock) of one instance, in a fixed tree order.  Source
// either a per-clump array (AoS, width w) or, for the tangent check, the dot product of the
// adjoint seed with the dual part of the final state (w = 1).
__global__ __launch_bounds__(kObsBlock) void k_win_partial(
    const double *__restrict__ vals, int w, const double *__restrict__ Sb,
    const Dual *__restrict__ stD, const int64_t *__restrict__ inv, int64_t begin, int64_t count,
    int64_t cap, double *__restrict__ partial) {
  __shared__ double red[kObsBlock];
  const int64_t o = (int64_t)blockIdx.x * kObsBlock + threadIdx.x;
  const int64_t c = o < count ? inv[begin + o] : -1;
  if (vals) {
    for (int k = 0; k < w; k++) {
      const double r = tree_sum(red, c >= 0 ? vals[c * w + k] : 0.0);
      if (threadIdx.x == 0) partial[blockIdx.x * w + k] = r;
    }
  } else {
    double d = 0.0;
    if (c >= 0)
      for (int f = 0; f < S_NFIELDS; f++) d += Sb[f * cap + c] * stD[f * cap + c].d;
    const double r = tree_sum(red, d);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
  }
}

// Sums the block partials in block order (one thread per component).
__global__ void k_win_partial_final(const double *__restrict__ partial, int64_t nblocks, int w,
                                    double *__restrict__ out) {
  const int k = threadIdx.x;
  if (blockIdx.x != 0 || k >= w) return;
  double acc = 0.0;
  for (int64_t b = 0; b < nblocks; b++) acc += partial[b * w + k];
  out[k] = acc;
}

// Window start: copies the state, adds the velocity perturbation (reference order:
// dv[(begin + o) * 3 + k]) and, for the tangent forward, lifts it to dual numbers with the
// velocity tangent tdv.
__global__ void k_win_start(const double *__restrict__ st, double *__restrict__ w0,
                            const int64_t *__restrict__ inv, const double *__restrict__ dv,
                            const double *__restrict__ tdv, Dual *__restrict__ stD, int64_t n,
                            int64_t cap) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  const int64_t c = inv[o];
  for (int f = 0; f < S_NFIELDS; f++) {
    double v = st[f * cap + c];
    double t = 0.0;
    if (f >= S_VX && f <= S_VZ) {
      if (dv) v += dv[3 * o + (f - S_VX)];
      if (tdv) t = tdv[3 * o + (f - S_VX)];
    }
    w0[f * cap + c] = v;
    if (stD) stD[f * cap + c] = mkd(v, t);
  }
}

// Exports per clump, in reference order: the state (24 doubles, StateField order) and/or
// the start-velocity adjoint (3 doubles).
__global__ void k_win_export(const double *__restrict__ st, const double *__restrict__ Sb,
                             const int64_t *__restrict__ inv, double *__restrict__ state_out,
                             double *__restrict__ dv_grad, int64_t n, int64_t cap) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  const int64_t c = inv[o];
  if (state_out)
    for (int f = 0; f < S_NFIELDS; f++) state_out[o * S_NFIELDS + f] = st[f * cap + c];
  if (dv_grad)
    for (int k = 0; k < 3; k++) dv_grad[3 * o + k] = Sb[(S_VX + k) * cap + c];
}

// Largest value of a per-clump array (non-negative doubles) as an order-preserving integer.
constexpr int kWinMaxBlock = 256;
constexpr unsigned kWinMaxGrid = 512;

// Grid-stride over v, one atomic per workgroup (launched with kWinMaxBlock threads and at
// most kWinMaxGrid workgroups, so the atomics on *out stay few).
__global__ __launch_bounds__(kWinMaxBlock) void k_win_max(const double *__restrict__ v,
                                                          int64_t n, unsigned long long *out) {
  __shared__ double red[kWinMaxBlock / 64];
  double m = 0.0;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x)
    m = v[c] > m ? v[c] : m;
  for (int off = 32; off > 0; off >>= 1) {
    const double q = __shfl_xor(m, off);
    m = q > m ? q : m;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kWinMaxBlock / 64; w++) m = red[w] > m ? red[w] : m;
    if (m > 0.0) diag_max(out, (unsigned long long)__double_as_longlong(m));
  }
}

}  // namespace avf
