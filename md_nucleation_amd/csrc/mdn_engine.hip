// mdn_engine.hip -- device engine and C ABI of libmdn (gfx950).
//
// Two execution paths over the same replica data and the same per-term code
// (mdn_terms.h):
//  * per-call path: one replica, one launch per reference call
//    (velocity_calculation, update_position, check_pairlist/make_pair,
//    Potential::compute), atoms in global memory, FP64 atomics.  This is the
//    plain drop-in for MD::verlet; it also carries single replicas too large
//    for one workgroup's LDS (tiled vv kernel below).
//  * batched path: every replica in one persistent launch, one workgroup per
//    replica, its in-domain atoms resident in LDS across steps; the pair-list
//    rebuild runs inside the same kernel when the replica's own loopPair
//    exceeds loop_update.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "mdn.h"
#include "mdn_device.h"
#include "mdn_internal.h"
#include "mdn_terms.h"

namespace mdn {

static thread_local std::string g_err;
int set_error(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

// ======================= per-call kernels ================================
constexpr int kBlk = 256;

// local in-domain atom a -> global atom index (GStore order)
__device__ inline int gatom(const DevReplica &R, const LView &L, int a) {
  if (a < L.n_ion) return a;
  if (a < L.gbase) {
    int k = (a - L.vbase) / L.na_v, i = a - L.vbase - k * L.na_v;
    return R.off_vap + (R.in_vap[k] - 1 - R.n_gas) * L.na_v + i;
  }
  int k = (a - L.gbase) / L.na_g, i = a - L.gbase - k * L.na_g;
  return R.off_gas + (R.in_gas[k] - 1) * L.na_g + i;
}
__device__ inline double gmass(const DevReplica &R, const LView &L, int a) {
  if (a < L.n_ion) return R.sp[0].mass[a];
  if (a < L.gbase) return R.sp[2].mass[(a - L.vbase) % L.na_v];
  return R.sp[1].mass[(a - L.gbase) % L.na_g];
}

// MD::velocity_calculation, verlet.cpp:47-65 (region atoms only).  At the
// start of an energy step it also clears the energy buckets (Uzero,
// run.cpp:68-71); at the end of a step it counts in-domain atom-steps and,
// under the Nose-Hoover thermostat, scales the ion's velocities by
// exp(-zeta dt / 2) after the kick (MD::nosehoover_ion, verlet.cpp:36).
__global__ void k_kick(const DevReplica *__restrict__ Rp, int zeroU, int count, int nh) {
  const DevReplica &R = *Rp;
  LView L = view_from_dyn(R);
  int t = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
  if (t == 0 && zeroU)
    for (int k = 0; k < U_N; k++) R.dyn->U[k] = 0;
  if (t == 0 && count) { R.dyn->atom_steps += L.n_in; R.dyn->step++; }
  const double C = 0.5 * R.dt * kKickUnit;
  const double fac = nh ? nh_factor(R.dyn->zeta_ion, R.dt) : 1.0;
  for (int a = t; a < L.n_in; a += nt) {
    int g = gatom(R, L, a);
    double c2 = C / gmass(R, L, a);
    double *v = R.atom_v + 3 * (size_t)g, *f = R.atom_f + 3 * (size_t)g;
    v[0] += f[0] * c2;
    v[1] += f[1] * c2;
    v[2] += f[2] * c2;
    if (nh && a < L.n_ion) { v[0] *= fac; v[1] *= fac; v[2] *= fac; }
  }
}

// sum of m v^2 over the ion atoms, returned to every thread of the (single)
// workgroup; s_red holds >= 16 doubles
__device__ inline double ion_mv2(const DevReplica &R, double *s_red) {
  const DevSpecies &I = R.sp[0];
  int tid = threadIdx.x, nt = blockDim.x;
  double k = 0;
  for (int i = tid; i < I.natoms; i += nt) {
    const double *v = R.atom_v + 3 * (size_t)i;
    double m = I.mass[i];
    k += v[0] * v[0] * m + v[1] * v[1] * m + v[2] * v[2] * m;
  }
  k = wave_sum(k);
  if ((tid & 63) == 0) s_red[tid >> 6] = k;
  __syncthreads();
  double s = 0;
  for (int w = 0; w < (nt + 63) / 64; w++) s += s_red[w];
  return s;
}

// MD::velocity_scaling, nvt.cpp:5-16: the ion's velocities scaled to 300 K
// right after the first half-kick (verlet.cpp:22); one workgroup
__global__ void __launch_bounds__(1024) k_ion_scale(const DevReplica *__restrict__ Rp) {
  const DevReplica &R = *Rp;
  __shared__ double s_red[16];
  const DevSpecies &I = R.sp[0];
  double Tion = ion_temperature(ion_mv2(R, s_red), I.natoms);
  double ratio = sqrt(300.0 / Tion);
  for (int i = threadIdx.x; i < 3 * I.natoms; i += blockDim.x) R.atom_v[i] *= ratio;
}

// drift + force clear, verlet.cpp:76-88 / shake.cpp:6-24
__global__ void k_drift(const DevReplica *__restrict__ Rp, int save_old) {
  const DevReplica &R = *Rp;
  LView L = view_from_dyn(R);
  int t = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
  for (int a = t; a < L.n_in; a += nt) {
    size_t g = 3 * (size_t)gatom(R, L, a);
    double *x = R.atom_x + g, *v = R.atom_v + g, *f = R.atom_f + g;
    if (save_old) { R.atom_xo[g] = x[0]; R.atom_xo[g + 1] = x[1]; R.atom_xo[g + 2] = x[2]; }
    x[0] += v[0] * R.dt;
    x[1] += v[1] * R.dt;
    x[2] += v[2] * R.dt;
    f[0] = f[1] = f[2] = 0.0;
  }
}

// SHAKE per in-domain molecule with rigid pairs, shake.cpp:25-37
// one thread per in-domain molecule; a molecule of up to kShakeLdsAtoms
// atoms iterates on a copy of its positions and old positions in LDS (a
// slot per thread, interleaved over the block's threads), so the sweeps'
// dependent reads and writes do not go through the cache hierarchy;
// larger molecules iterate in place in global memory.  Either way the
// arithmetic is shake_molecule's, so the positions are the same.
constexpr int kShakeBlk = 64, kShakeLdsAtoms = 12;
__global__ void __launch_bounds__(kShakeBlk) k_shake(const DevReplica *__restrict__ Rp, int skip_vapor) {
  const DevReplica &R = *Rp;
  __shared__ double s_x[kShakeLdsAtoms * 3 * kShakeBlk], s_xo[kShakeLdsAtoms * 3 * kShakeBlk];
  LView L = view_from_dyn(R);
  const int tid = threadIdx.x;
  int t = blockIdx.x * blockDim.x + tid, nt = gridDim.x * blockDim.x;
  int nm = 1 + L.nv + L.ng;
  // the largest iteration count of the thread's molecules, merged over the
  // wave before the one atomic per wave (thousands of same-address atomics
  // would serialise at the cache)
  int itmax = 0;
  for (int m = t; m < nm; m += nt) {
    int sp, o;
    if (m == 0) { sp = 0; o = 0; }
    else if (m <= L.nv) { sp = 2; o = R.off_vap + (R.in_vap[m - 1] - 1 - R.n_gas) * L.na_v; }
    else { sp = 1; o = R.off_gas + (R.in_gas[m - 1 - L.nv] - 1) * L.na_g; }
    const DevSpecies &S = R.sp[sp];
    if (!S.nrigid || (sp == 2 && skip_vapor)) continue;
    double *x = R.atom_x + 3 * (size_t)o, *v = R.atom_v + 3 * (size_t)o;
    const double *xo = R.atom_xo + 3 * (size_t)o;
    auto V = [&](int i, int q) -> double & { return v[3 * i + q]; };
    auto finish = [&](auto X, auto XO, int it) {
      if (it < 0) {
        atomicAdd(&R.dyn->shake_fail, 1ull);
        itmax = max(itmax, -it);
        shake_undo(S, X, XO, V, R.dt);
        return;
      }
      itmax = max(itmax, it);
      if (R.shake == 1) shake_velocity(S, X, XO, V, R.dt);
    };
    if (S.natoms <= kShakeLdsAtoms) {
      auto XL = [&](int i, int q) -> double & { return s_x[(3 * i + q) * kShakeBlk + tid]; };
      auto XOL = [&](int i, int q) -> double { return s_xo[(3 * i + q) * kShakeBlk + tid]; };
      for (int k = 0; k < 3 * S.natoms; k++) {
        s_x[k * kShakeBlk + tid] = x[k];
        s_xo[k * kShakeBlk + tid] = xo[k];
      }
      finish(XL, XOL, shake_molecule(R, S, XL, XOL, R.shake_max_iter));
      for (int k = 0; k < 3 * S.natoms; k++) x[k] = s_x[k * kShakeBlk + tid];
    } else {
      auto X = [&](int i, int q) -> double & { return x[3 * i + q]; };
      auto XO = [&](int i, int q) -> double { return xo[3 * i + q]; };
      finish(X, XO, shake_molecule(R, S, X, XO, R.shake_max_iter));
    }
  }
  // every lane of the block reaches this point (no early return above)
  for (int k = 32; k > 0; k >>= 1) itmax = max(itmax, __shfl_xor(itmax, k));
  if ((tid & 63) == 0 && itmax > 0) atomicMax(&R.dyn->shake_iters_max, itmax);
}

// SHAKE of the in-domain vapor molecules of a species whose constraints fall
// into star groups (DevSpecies::shake_star: each group joins one centre atom
// to at most three satellites; methanol under SHAKE H: C-H3 and O-H): one
// lane per (molecule, group), a molecule's G lanes adjacent within a quad,
// each group iterating with its atoms in registers (shake_group_star) and the
// lanes agreeing on convergence after every sweep.  The groups move disjoint
// atoms, so the positions are those of k_shake's sweeps over the molecule,
// to the bit; the longest chain of dependent updates per sweep is the
// largest group's instead of the molecule's.  The ion and gas molecules stay
// with k_shake.
constexpr int kShakeGrpBlk = 256, kShakeGrpCap = 16;
__global__ void __launch_bounds__(kShakeGrpBlk) k_shake_groups(const DevReplica *__restrict__ Rp) {
  const DevReplica &R = *Rp;
  __shared__ int s_n[2], s_ij[2 * kShakeGrpCap * 2], s_ng[2], s_gs[10], s_star[2];
  __shared__ double s_c[2 * kShakeGrpCap * 4];
  __shared__ unsigned long long s_cons[2], s_gm[8];
  ShakeLds T;
  T.n = s_n; T.ij = s_ij; T.c = s_c; T.cons = s_cons; T.ng = s_ng; T.gs = s_gs; T.gm = s_gm; T.star = s_star;
  T.cap = kShakeGrpCap;
  const DevSpecies &S = R.sp[2];
  if (threadIdx.x == 0) shake_lds_fill(R, S, T, 1);
  __syncthreads();
  const LView L = view_from_dyn(R);
  const int G = T.ng[1], gsh = G >= 4 ? 2 : G >= 2 ? 1 : 0, W = 1 << gsh;
  const int nslot = L.nv << gsh;
  const int t = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
  int itmax = 0;
  // slots advance by nt (a multiple of 4): a molecule's lanes stay in one
  // quad and are all active or all idle
  for (int slot = t; slot < nslot; slot += nt) {
    const int m = slot >> gsh, g = slot & (W - 1);
    const size_t o = (size_t)R.off_vap + (size_t)(R.in_vap[m] - 1 - R.n_gas) * L.na_v;
    double *x = R.atom_x + 3 * o, *v = R.atom_v + 3 * o;
    const double *xo = R.atom_xo + 3 * o;
    auto X = [&](int i, int q) -> double & { return x[3 * i + q]; };
    auto XO = [&](int i, int q) -> double { return xo[3 * i + q]; };
    auto V = [&](int i, int q) -> double & { return v[3 * i + q]; };
    const int it = shake_group_lds(R, T, 1, g, W, X, XO, R.shake_max_iter);
    if (it < 0) {
      // every lane of the molecule sees the failure; the first undoes it
      itmax = max(itmax, -it);
      if (g == 0) {
        atomicAdd(&R.dyn->shake_fail, 1ull);
        shake_undo(S, X, XO, V, R.dt);
      }
      continue;
    }
    itmax = max(itmax, it);
    if (R.shake == 1) {
      const unsigned long long gm = T.gm[4 + g];
      for (int a = 0; a < S.natoms; a++)
        if ((gm >> a) & 1ull)
          for (int q = 0; q < 3; q++) V(a, q) = (X(a, q) - XO(a, q)) / R.dt;
    }
  }
  for (int k = 32; k > 0; k >>= 1) itmax = max(itmax, __shfl_xor(itmax, k));
  if ((threadIdx.x & 63) == 0 && itmax > 0) atomicMax(&R.dyn->shake_iters_max, itmax);
}

// Observer::computeProps (observer.cpp:5-36), one workgroup per replica:
// out[8 r ..] = Kion, K_g, K_v, Tion, T_g, T_v.  Gas and vapor use every
// molecule's COM velocity (as of its last rebuild, or its draw when out of
// the domain) and declared molecule mass; the ion its atom velocities.
__global__ void __launch_bounds__(256) k_observe(const DevReplica *__restrict__ reps, int r0, double *out) {
  const DevReplica &R = reps[r0 + blockIdx.x];
  __shared__ double s_red[3][4];
  const int tid = threadIdx.x, nt = blockDim.x;
  double k[3] = {0, 0, 0};
  const DevSpecies &I = R.sp[0];
  for (int i = tid; i < I.natoms; i += nt) {
    const double *v = R.atom_v + 3 * (size_t)i;
    double m = I.mass[i];
    k[0] += v[0] * v[0] * m + v[1] * v[1] * m + v[2] * v[2] * m;
  }
  for (int i = 1 + tid; i < R.n_mol; i += nt) {
    const double *v = R.mol_v + 3 * (size_t)i;
    const int sp = i <= R.n_gas ? 1 : 2;
    double m = R.sp[sp].molmass;
    k[sp] += v[0] * v[0] * m + v[1] * v[1] * m + v[2] * v[2] * m;
  }
  for (int q = 0; q < 3; q++) {
    double sv = wave_sum(k[q]);
    if ((tid & 63) == 0) s_red[q][tid >> 6] = sv;
  }
  __syncthreads();
  if (tid < 3) {
    double sv = 0;
    for (int w = 0; w < nt / 64; w++) sv += s_red[tid][w];
    const double K = sv * (0.5 * kRealToKcal);
    const int n = tid == 0 ? I.natoms : (tid == 1 ? R.n_gas : R.n_vap);
    double *o = out + 8 * (size_t)blockIdx.x;
    o[tid] = K;
    o[3 + tid] = K / (double)n * ((1 / kKbReal) / 1.5);
  }
}

// ion COM + blend weights, verlet.cpp:90-134; with nh set, also the
// Nose-Hoover friction update that follows update_position
// (MD::nosehoover_zeta, verlet.cpp:27).  Every workgroup sums the ion COM
// itself (the same sums in the same order, so the same value) and weighs
// its share of the in-domain atoms; workgroup 0 writes the COM and zeta.
__global__ void __launch_bounds__(1024) k_com_weights(const DevReplica *__restrict__ Rp, int nh) {
  const DevReplica &R = *Rp;
  __shared__ double s_red[7][16];
  __shared__ double s_com[6];
  LView L = view_from_dyn(R);
  int tid = threadIdx.x, nt = blockDim.x;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  const DevSpecies &I = R.sp[0];
  for (int i = tid; i < I.natoms; i += nt) {
    double m = I.mass[i];
    const double *x = R.atom_x + 3 * (size_t)i, *v = R.atom_v + 3 * (size_t)i;
    acc[0] += x[0] * m; acc[1] += x[1] * m; acc[2] += x[2] * m;
    acc[3] += v[0] * m; acc[4] += v[1] * m; acc[5] += v[2] * m;
    acc[6] += v[0] * v[0] * m + v[1] * v[1] * m + v[2] * v[2] * m;
  }
  for (int k = 0; k < 7; k++) {
    double s = wave_sum(acc[k]);
    if ((tid & 63) == 0) s_red[k][tid >> 6] = s;
  }
  __syncthreads();
  if (tid < 7) {
    double s = 0;
    for (int w = 0; w < (nt + 63) / 64; w++) s += s_red[tid][w];
    if (blockIdx.x != 0) {
      if (tid < 6) s_com[tid] = s * (1 / I.molmass);
    } else if (tid == 6) {
      if (nh) R.dyn->zeta_ion = nh_zeta_step(R.dyn->zeta_ion, ion_temperature(s, I.natoms), R.dyn->Tnh_ion,
                                             I.natoms, R.dt);
    } else {
      double c = s * (1 / I.molmass);
      s_com[tid] = c;
      if (tid < 3) R.mol_x[tid] = c; else R.mol_v[tid - 3] = c;
    }
  }
  __syncthreads();
  for (int a = L.n_ion + blockIdx.x * nt + tid; a < L.n_in; a += nt * gridDim.x) {
    int g = gatom(R, L, a);
    const double *x = R.atom_x + 3 * (size_t)g;
    double dx = x[0] - s_com[0], dy = x[1] - s_com[1], dz = x[2] - s_com[2];
    R.atom_w[g] = blend_weight(dx * dx + dy * dy + dz * dz);
  }
}

// MD::check_pairlist + MD::make_pair (pairlist.cpp:16-63), one workgroup
__global__ void __launch_bounds__(1024) k_rebuild(const DevReplica *__restrict__ Rp, int force) {
  const DevReplica &R = *Rp;
  __shared__ int s_int[64];
  __shared__ double s_dbl[16];
  __shared__ double s_big[3 * 1024];
  __shared__ int s_do;
  if (threadIdx.x == 0) {
    int go = force;
    if (!force) {
      int lp = ++R.dyn->loopPair;
      go = lp > R.dyn->loop_update;
    }
    s_do = go;
  }
  __syncthreads();
  if (s_do) rebuild_wg(R, s_int, s_dbl, s_big, 3 * 1024);
}

// wave-reduced energy flush
__device__ inline void flush_energy(const DevReplica &R, const double *e) {
  for (int k = 0; k < U_N; k++) {
    double s = wave_sum(e[k]);
    if ((threadIdx.x & 63) == 0 && s != 0.0) atomicAdd(&R.dyn->U[k], s);
  }
}

template <typename Rl>
__global__ void __launch_bounds__(kBlk) k_terms(const DevReplica *__restrict__ Rp, uint32_t bits, int eflag) {
  const DevReplica &R = *Rp;
  GStore s;
  s.R = Rp;
  s.L = view_from_dyn(R);
  int t = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
  double e[U_N] = {0, 0, 0, 0, 0, 0, 0, 0};
  all_terms<Rl>(R, s, bits, t, nt, eflag != 0, e);
  if (eflag) flush_energy(R, e);
}


// every lane takes the value of lane (l + 1) mod 64 (DPP wave_rol:1; all 64
// lanes must be active)
__device__ inline double wave_rol1(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x134, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x134, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <typename Rl> struct Pack4;
template <> struct Pack4<float> { typedef float4 T; };
template <> struct Pack4<double> { typedef double4 T; };

// vv for in-domain vapor sets of up to a few thousand atoms: i-centric, j
// tiles staged through LDS, no atomics on j; grid.y splits the j range so
// that a launch fills the chip even for a few hundred atoms (the Newton-3
// kernel below, one wave per 64 x 64 block pair, would leave such a set
// with a handful of long serial chains).  Each unordered pair is evaluated
// from both sides, so its energy is halved.  The partner loop has no
// branches: pairs of one molecule (the atom itself included) take r^2 = 1
// and weight 0, and a tile is padded to a multiple of four partners with
// far-away, zero-weight, zero-charge entries, so four partners' LDS reads
// issue together; a padding partner's energy is masked off.  Per type pair
// the table holds c0, c1, c0 / 12 and c1 / 6 in pair precision (the same
// values as dividing per pair).

constexpr int kVVTile = 256;
// workgroup (bx, by) of a gx x gy grid
template <typename Rl, bool EF>
__device__ void vv_tiled(const DevReplica *__restrict__ Rp, int tv0, int ntv, int bx, int by, int gy) {
  const DevReplica &R = *Rp;
  __shared__ double sx[kVVTile * 3];
  __shared__ Rl sq[kVVTile], sw[kVVTile];
  __shared__ int st[kVVTile], sm[kVVTile];
  __shared__ typename Pack4<Rl>::T sc[64];
  LView L = view_from_dyn(R);
  const int nva = L.nv * L.na_v;
  // vapor-vapor coefficient table (vapor types are contiguous)
  for (int k = threadIdx.x; k < ntv * ntv; k += blockDim.x) {
    int a = k / ntv, b = k - a * ntv;
    const double *p = R.pair_coeff + 2 * ((size_t)(tv0 + a) * R.n_types + tv0 + b);
    const Rl c0 = (Rl)p[0], c1 = (Rl)p[1];
    sc[k] = {c0, c1, c0 / Rl(12), c1 / Rl(6)};
  }
  const int i = bx * blockDim.x + threadIdx.x;
  const bool act = i < nva;
  double xi[3] = {0, 0, 0}, qi = 0, wi = 0;
  int ti = 0, mi = -1, gi = 0;
  const DevSpecies &V = R.sp[2];
  if (act) {
    int k = i / L.na_v, ii = i - k * L.na_v;
    gi = R.off_vap + (R.in_vap[k] - 1 - R.n_gas) * L.na_v + ii;
    xi[0] = R.atom_x[3 * (size_t)gi]; xi[1] = R.atom_x[3 * (size_t)gi + 1]; xi[2] = R.atom_x[3 * (size_t)gi + 2];
    qi = R.qqrd2e * V.charge[ii];
    wi = R.atom_w[gi];
    ti = V.type[ii] - tv0;
    mi = k;
  }
  const int trow = ti * ntv;
  // charge (times qqrd2e) and weight products in pair precision
  const Rl qr = (Rl)qi, wr = (Rl)wi;
  int per = (nva + gy - 1) / gy;
  int j0 = by * per, j1 = min(nva, j0 + per);
  double fx = 0, fy = 0, fz = 0, e = 0;
  // one partner b; valid: b is a real entry of the tile (uniform)
  auto pair = [&](int b, bool valid) {
    const double dx = xi[0] - sx[3 * b], dy = xi[1] - sx[3 * b + 1], dz = xi[2] - sx[3 * b + 2];
    const bool on = sm[b] != mi;
    const Rl r2 = (Rl)(dx * dx + dy * dy + dz * dz);
    const typename Pack4<Rl>::T cc = sc[trow + st[b]];
    const Rl qq = qr * sq[b], w = wr * sw[b];
    Rl r2inv, rinv;
    inv_r<Rl>(on ? r2 : (Rl)1, r2inv, rinv);
    const Rl r6inv = r2inv * r2inv * r2inv;
    const Rl flj = r6inv * (cc.x * r6inv - cc.y);
    const Rl fc = qq * rinv;
    if (EF) {
      const double ep = (double)(r6inv * (cc.z * r6inv - cc.w)) + (double)fc;
      e += on && valid ? ep : 0.0;
    }
    // the pair-precision scalar times the FP64 displacement
    const double fpd = (double)((flj + fc) * r2inv * (on ? w : (Rl)0));
    fx = fma(fpd, dx, fx); fy = fma(fpd, dy, fy); fz = fma(fpd, dz, fz);
  };
  for (int jt = j0; jt < j1; jt += kVVTile) {
    __syncthreads();
    const int nj = min(kVVTile, j1 - jt), nj4 = (nj + 3) & ~3;
    const int b = threadIdx.x;
    if (b < nj) {
      const int j = jt + b;
      int k = j / L.na_v, jj = j - k * L.na_v;
      int g = R.off_vap + (R.in_vap[k] - 1 - R.n_gas) * L.na_v + jj;
      sx[3 * b] = R.atom_x[3 * (size_t)g];
      sx[3 * b + 1] = R.atom_x[3 * (size_t)g + 1];
      sx[3 * b + 2] = R.atom_x[3 * (size_t)g + 2];
      sq[b] = (Rl)V.charge[jj];
      sw[b] = (Rl)R.atom_w[g];
      st[b] = V.type[jj] - tv0;
      sm[b] = k;
    } else if (b < nj4) {
      sx[3 * b] = 1e6; sx[3 * b + 1] = 1e6; sx[3 * b + 2] = 1e6;
      sq[b] = 0; sw[b] = 0; st[b] = 0; sm[b] = -2;
    }
    __syncthreads();
    if (act) {
      for (int b4 = 0; b4 < nj4; b4 += 4) {
        pair(b4, true);
        pair(b4 + 1, b4 + 1 < nj);
        pair(b4 + 2, b4 + 2 < nj);
        pair(b4 + 3, b4 + 3 < nj);
      }
    }
  }
  if (act) {
    atomicAdd(&R.atom_f[3 * (size_t)gi], fx);
    atomicAdd(&R.atom_f[3 * (size_t)gi + 1], fy);
    atomicAdd(&R.atom_f[3 * (size_t)gi + 2], fz);
  }
  if (EF) {
    double s = wave_sum(0.5 * e);
    if ((threadIdx.x & 63) == 0 && s != 0.0) atomicAdd(&R.dyn->U[U_VV], s);
  }
}

template <typename Rl, bool EF>
__global__ void __launch_bounds__(kVVTile) k_vv_tiled(const DevReplica *__restrict__ Rp, int tv0, int ntv) {
  vv_tiled<Rl, EF>(Rp, tv0, ntv, blockIdx.x, blockIdx.y, gridDim.y);
}

// k_terms on workgroups [0, kTermsGrid) and the tiled vv (a gx x gy grid,
// row-major) on the rest, in one launch: both only add into the force
// array, so they run side by side rather than one after the other
constexpr int kTermsGrid = 512;
template <typename Rl, bool EF>
__global__ void __launch_bounds__(kBlk) k_terms_vv(const DevReplica *__restrict__ Rp, uint32_t bits, int tv0, int ntv,
                                                   int gx, int gy) {
  if ((int)blockIdx.x < kTermsGrid) {
    const DevReplica &R = *Rp;
    GStore s;
    s.R = Rp;
    s.L = view_from_dyn(R);
    const int t = blockIdx.x * blockDim.x + threadIdx.x, nt = kTermsGrid * kBlk;
    double e[U_N] = {0, 0, 0, 0, 0, 0, 0, 0};
    all_terms<Rl>(R, s, bits, t, nt, EF, e);
    if (EF) flush_energy(R, e);
  } else {
    const int b = blockIdx.x - kTermsGrid;
    vv_tiled<Rl, EF>(Rp, tv0, ntv, b % gx, b / gx, gy);
  }
}

// vv (VaporVapor.cpp:5-45) for large in-domain vapor sets (e.g. a
// 2,000-methanol cluster), each unordered pair evaluated once (Newton's
// third law).  The in-domain vapor atoms are cut into 64-atom blocks and
// those into super-blocks of four; a 256-thread workgroup takes one
// super-block pair (SI <= SJ).  Its four waves own the four i-blocks of SI
// (lane l of wave w: atom 256 SI + 64 w + l, held in registers) and visit
// the j-blocks of SJ in the order w, w + 1, w + 2, w + 3 (mod 4), so that
// at any stage the waves work on different j-blocks.  Within a block pair
// lane l meets, at step k, partner slot (l + k) mod 64: F_i accumulates in
// FP64 registers, F_j in FP64 registers that travel with their partner (a
// wave rotate after every step hands lane l the accumulator of lane l + 1,
// whose partner it meets next) and are added to the workgroup's LDS F_j
// array when the block pair ends.  At the end every F_i and F_j reaches
// atom_f by one FP64 atomic per component.  An off-diagonal block pair runs
// steps k and k + 32 together (two chains), k = 0..31; a block with itself
// runs k = 1..32, k = 32 on lanes 0..31 only.  In a diagonal super-block
// pair (SI = SJ) wave w takes the block pairs (w, w), (w, w + 1 mod 4) and,
// for w < 2, (w, w + 2), which covers each unordered pair once.  A vapor
// molecule's atoms are contiguous and at most 64, so they lie in at most
// two adjacent blocks: the same-molecule test is compiled only into block
// pairs of adjacent or equal blocks.  Padding atoms past the vapor range sit
// 1e6 A away at distinct places, with zero charge and weight: their forces
// are exactly zero and their energies are masked.  Per type pair the table
// holds c0, c1, c0 / 12 and c1 / 6 in pair precision.
constexpr int kVVSuper = 256;

// an atom's pair data: the i side carries qqrd2e q, the j side q
template <typename Rl>
struct alignas(16) VVp {
  Rl q, w;
  int t, m;  // type - first vapor type; in-domain molecule (-1: padding)
};

template <typename Rl>
struct VVI {
  double x0, x1, x2;
  VVp<Rl> p;
};

template <typename Rl>
struct VVLds {
  double *x, *y, *z;         // [256] partner positions
  VVp<Rl> *p;                // [256]
  double *fx, *fy, *fz;      // [256] partner force sums
  typename Pack4<Rl>::T *sc; // [ntv * ntv]
};

// atom a (< 256 nsb) of the in-domain vapor range, or padding past nva
template <typename Rl>
__device__ inline void vv_atom(const DevReplica &R, const LView &L, int nva, int tv0, int a, bool iside, double &x0,
                               double &x1, double &x2, VVp<Rl> &p) {
  if (a < nva) {
    const DevSpecies &V = R.sp[2];
    const int k = a / L.na_v, ii = a - k * L.na_v;
    const size_t g = R.off_vap + (size_t)(R.in_vap[k] - 1 - R.n_gas) * L.na_v + ii;
    x0 = R.atom_x[3 * g]; x1 = R.atom_x[3 * g + 1]; x2 = R.atom_x[3 * g + 2];
    p.q = (Rl)(iside ? R.qqrd2e * V.charge[ii] : V.charge[ii]);
    p.w = (Rl)R.atom_w[g];
    p.t = V.type[ii] - tv0;
    p.m = k;
  } else {
    x0 = 1e6 + 1e3 * (double)a; x1 = 1e6; x2 = 1e6;
    p.q = 0; p.w = 0; p.t = 0; p.m = -1;
  }
}

__device__ inline size_t vv_global(const DevReplica &R, const LView &L, int a) {
  const int k = a / L.na_v, ii = a - k * L.na_v;
  return R.off_vap + (size_t)(R.in_vap[k] - 1 - R.n_gas) * L.na_v + ii;
}

// one pair: returns s with the force on i = s (x_i - x_j), on j the
// opposite; pass = 0 masks the pair
template <typename Rl, bool EF, bool NEAR>
__device__ inline double vv_n3_pair(const VVI<Rl> &c, const VVLds<Rl> &S, int j, int ntv, uint32_t pass, double &dx,
                                    double &dy, double &dz, double &e) {
  const VVp<Rl> pj = S.p[j];
  dx = c.x0 - S.x[j];
  dy = c.x1 - S.y[j];
  dz = c.x2 - S.z[j];
  const uint32_t on = NEAR ? pass & (pj.m != c.p.m ? 1u : 0u) : 1u;
  const Rl r2 = (Rl)(dx * dx + dy * dy + dz * dz);
  const Rl rsq = NEAR ? (on ? r2 : (Rl)1) : r2;
  const typename Pack4<Rl>::T cc = S.sc[c.p.t * ntv + pj.t];
  const Rl qq = c.p.q * pj.q, w = c.p.w * pj.w;
  Rl r2inv, rinv;
  inv_r<Rl>(rsq, r2inv, rinv);
  const Rl r6inv = r2inv * r2inv * r2inv;
  const Rl flj = r6inv * (cc.x * r6inv - cc.y);
  const Rl fc = qq * rinv;
  if (EF) {
    const double ep = (double)(r6inv * (cc.z * r6inv - cc.w)) + (double)fc;
    e += on && pj.m >= 0 && c.p.m >= 0 ? ep : 0.0;
  }
  return (double)((flj + fc) * r2inv * (NEAR ? (on ? w : (Rl)0) : w));
}

// i-block against a different j-block (slots jb..jb+63): steps k and
// k + 32 together; F_j sums into the LDS array
template <typename Rl, bool EF, bool NEAR>
__device__ inline void vv_n3_off(const VVI<Rl> &c, const VVLds<Rl> &S, int jb, int lane, int ntv, double &fx,
                                 double &fy, double &fz, double &e) {
  double a0x = 0, a0y = 0, a0z = 0, a1x = 0, a1y = 0, a1z = 0;
#pragma unroll 2
  for (int t = 0; t < 32; t++) {
    double dx, dy, dz, ex, ey, ez;
    const double s = vv_n3_pair<Rl, EF, NEAR>(c, S, jb + ((lane + t) & 63), ntv, 1u, dx, dy, dz, e);
    const double u = vv_n3_pair<Rl, EF, NEAR>(c, S, jb + ((lane + t + 32) & 63), ntv, 1u, ex, ey, ez, e);
    fx = fma(s, dx, fx); fy = fma(s, dy, fy); fz = fma(s, dz, fz);
    fx = fma(u, ex, fx); fy = fma(u, ey, fy); fz = fma(u, ez, fz);
    a0x = wave_rol1(fma(-s, dx, a0x));
    a0y = wave_rol1(fma(-s, dy, a0y));
    a0z = wave_rol1(fma(-s, dz, a0z));
    a1x = wave_rol1(fma(-u, ex, a1x));
    a1y = wave_rol1(fma(-u, ey, a1y));
    a1z = wave_rol1(fma(-u, ez, a1z));
  }
  // after 32 rotations lane l holds chain 0's sum of slot l + 32 and chain
  // 1's of slot l
  const int s0 = jb + ((lane + 32) & 63), s1 = jb + lane;
  atomicAdd(&S.fx[s0], a0x); atomicAdd(&S.fy[s0], a0y); atomicAdd(&S.fz[s0], a0z);
  atomicAdd(&S.fx[s1], a1x); atomicAdd(&S.fy[s1], a1y); atomicAdd(&S.fz[s1], a1z);
}

// a block with itself (its atoms also in slots jb..jb+63): k = 1..32
template <typename Rl, bool EF>
__device__ inline void vv_n3_diag(const VVI<Rl> &c, const VVLds<Rl> &S, int jb, int lane, int ntv, double &fx,
                                  double &fy, double &fz, double &e) {
  double ax = 0, ay = 0, az = 0;
#pragma unroll 2
  for (int k = 1; k <= 32; k++) {
    double dx, dy, dz;
    const uint32_t pass = k < 32 || lane < 32 ? 1u : 0u;
    const double s = vv_n3_pair<Rl, EF, true>(c, S, jb + ((lane + k) & 63), ntv, pass, dx, dy, dz, e);
    fx = fma(s, dx, fx); fy = fma(s, dy, fy); fz = fma(s, dz, fz);
    ax = wave_rol1(fma(-s, dx, ax));
    ay = wave_rol1(fma(-s, dy, ay));
    az = wave_rol1(fma(-s, dz, az));
  }
  // after 32 rotations lane l holds the sum of slot l + 33
  const int s0 = jb + ((lane + 33) & 63);
  atomicAdd(&S.fx[s0], ax); atomicAdd(&S.fy[s0], ay); atomicAdd(&S.fz[s0], az);
}

template <typename Rl, bool EF>
__global__ void __launch_bounds__(kVVSuper) k_vv_n3(const DevReplica *__restrict__ Rp, int tv0, int ntv) {
  const DevReplica &R = *Rp;
  __shared__ double sx[kVVSuper], sy[kVVSuper], sz[kVVSuper], sfx[kVVSuper], sfy[kVVSuper], sfz[kVVSuper];
  __shared__ VVp<Rl> sp[kVVSuper];
  __shared__ typename Pack4<Rl>::T sc[64];
  const LView L = view_from_dyn(R);
  const int nva = L.nv * L.na_v;
  const int nsb = (nva + kVVSuper - 1) / kVVSuper;
  // super-block pair of this workgroup: row-major over SI <= SJ
  int t = blockIdx.x, SI = 0;
  while (SI < nsb && t >= nsb - SI) { t -= nsb - SI; SI++; }
  if (SI >= nsb) return;  // the grid is sized for every vapor molecule in the domain
  const int SJ = SI + t;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int k = tid; k < ntv * ntv; k += kVVSuper) {
    const int a = k / ntv, b = k - a * ntv;
    const double *p = R.pair_coeff + 2 * ((size_t)(tv0 + a) * R.n_types + tv0 + b);
    const Rl c0 = (Rl)p[0], c1 = (Rl)p[1];
    sc[k] = {c0, c1, c0 / Rl(12), c1 / Rl(6)};
  }
  {
    VVp<Rl> p;
    double x0, x1, x2;
    vv_atom<Rl>(R, L, nva, tv0, kVVSuper * SJ + tid, false, x0, x1, x2, p);
    sx[tid] = x0; sy[tid] = x1; sz[tid] = x2; sp[tid] = p;
    sfx[tid] = 0; sfy[tid] = 0; sfz[tid] = 0;
  }
  VVI<Rl> c;
  const int ia = kVVSuper * SI + 64 * w + lane;
  vv_atom<Rl>(R, L, nva, tv0, ia, true, c.x0, c.x1, c.x2, c.p);
  __syncthreads();
  const VVLds<Rl> S{sx, sy, sz, sp, sfx, sfy, sfz, sc};
  double fx = 0, fy = 0, fz = 0, e = 0;
  if (SI == SJ) {
    vv_n3_diag<Rl, EF>(c, S, 64 * w, lane, ntv, fx, fy, fz, e);
    vv_n3_off<Rl, EF, true>(c, S, 64 * ((w + 1) & 3), lane, ntv, fx, fy, fz, e);
    if (w < 2) vv_n3_off<Rl, EF, true>(c, S, 64 * (w + 2), lane, ntv, fx, fy, fz, e);
  } else {
    for (int s = 0; s < 4; s++) {
      const int jw = (w + s) & 3;
      // blocks 4 SI + w and 4 SJ + jw are adjacent only for SJ = SI + 1,
      // w = 3, jw = 0
      if (SJ == SI + 1 && w == 3 && jw == 0) vv_n3_off<Rl, EF, true>(c, S, 64 * jw, lane, ntv, fx, fy, fz, e);
      else vv_n3_off<Rl, EF, false>(c, S, 64 * jw, lane, ntv, fx, fy, fz, e);
    }
  }
  if (ia < nva) {
    const size_t g = vv_global(R, L, ia);
    atomicAdd(&R.atom_f[3 * g], fx);
    atomicAdd(&R.atom_f[3 * g + 1], fy);
    atomicAdd(&R.atom_f[3 * g + 2], fz);
  }
  __syncthreads();
  const int ja = kVVSuper * SJ + tid;
  if (ja < nva) {
    const size_t g = vv_global(R, L, ja);
    atomicAdd(&R.atom_f[3 * g], sfx[tid]);
    atomicAdd(&R.atom_f[3 * g + 1], sfy[tid]);
    atomicAdd(&R.atom_f[3 * g + 2], sfz[tid]);
  }
  if (EF) {
    const double s = wave_sum(e);
    if (lane == 0 && s != 0.0) atomicAdd(&R.dyn->U[U_VV], s);
  }
}

// ======================= batched persistent kernel ========================
// LDS of one workgroup: per in-domain atom x, v, f (FP64), w, kick factor
// (FP64), charge and LJ mixing parameters (pair precision), type; old
// positions only for the ion + vapor atoms (the only constrained ones).
// per-atom pair parameters packed for one LDS read: charge, sqrt(eps),
// sigma, local vapor molecule id (-1: ion or gas) as a value


// per-atom LDS record: FP64 position and the pair-precision blend weight in
// one 32-byte slot (the dense pass finds both from one address)
template <typename Rl>
struct XRec {
  double x[3];
  Rl w;
};

template <typename Rl>
struct BatchLds {
  XRec<Rl> *xr;
  double *v, *f, *xo;
  double *kct;      // [kcap] kick factor 0.5 dt 4.184e-4 / m per species atom (ion | vapor | gas)
  // packed (q, sqrt(24 eps), sigma / 2, vapor molecule id or -1): pk per
  // atom of the ion + vapor range [cap_h], kpk per species atom [kcap]
  // (what gas atoms use)
  typename Pack4<Rl>::T *pk, *kpk;
  int16_t *kidx;    // per atom: index into kct
  uint32_t *allow;  // [acap] AMBER non-excluded ion-pair bitmask (rows of amber_words)
  double *red;      // [U_N + 6][16]
  double *sdbl;     // [16]
  double *kms;      // [kcap] species-atom masses
  ShakeLds shk;     // constraint tables of the ion and the vapor [2][scap]
  int *sint;        // [64]
};

template <typename Rl>
__device__ inline BatchLds<Rl> carve(char *base, int cap, int cap_h, int kcap, int acap, int scap) {
  BatchLds<Rl> B;
  double *d = reinterpret_cast<double *>(base);
  B.xr = reinterpret_cast<XRec<Rl> *>(d); d += 4 * cap;
  B.v = d; d += 3 * cap;
  B.f = d; d += 3 * cap;
  B.xo = d; d += 3 * cap_h;
  B.kct = d; d += kcap;
  B.red = d; d += (U_N + 6) * 16;
  B.sdbl = d; d += 16;
  B.kms = d; d += kcap;
  B.shk.c = d; d += 8 * (size_t)scap;
  B.shk.cons = reinterpret_cast<unsigned long long *>(d); d += 2;
  B.shk.gm = reinterpret_cast<unsigned long long *>(d); d += 8;
  B.shk.cap = scap;
  B.pk = reinterpret_cast<typename Pack4<Rl>::T *>(d);
  d += (4 * sizeof(Rl) * (size_t)cap_h) / sizeof(double);
  B.kpk = reinterpret_cast<typename Pack4<Rl>::T *>(d);
  d += (4 * sizeof(Rl) * (size_t)kcap + sizeof(double) - 1) / sizeof(double);
  int *ip = reinterpret_cast<int *>(d);
  B.sint = ip; ip += 64;
  B.shk.n = ip; ip += 2;
  B.shk.ij = ip; ip += 4 * (size_t)scap;
  B.shk.ng = ip; ip += 2;
  B.shk.gs = ip; ip += 10;
  B.shk.star = ip; ip += 2;
  B.allow = reinterpret_cast<uint32_t *>(ip); ip += acap;
  int16_t *hp = reinterpret_cast<int16_t *>(ip);
  B.kidx = hp; hp += cap;
  return B;
}

static size_t batch_lds_bytes(int cap, int cap_h, int kcap, int acap, int scap, size_t rl) {
  return sizeof(double) * ((size_t)10 * cap + 3 * (size_t)cap_h + 2 * (size_t)kcap + (U_N + 6) * 16 + 16 +
                           8 * (size_t)scap + 2 + 8) +
         sizeof(int) * (2 + 4 * (size_t)scap + 14) +
         ((4 * rl * (size_t)kcap + sizeof(double) - 1) / sizeof(double)) * sizeof(double) +
         rl * 4 * (size_t)cap_h + sizeof(int) * (64 + (size_t)acap) + sizeof(int16_t) * (size_t)cap;
}

// LDS store with Lorentz-Berthelot mixing from per-atom parameters
// (setCrossPotential.cpp:13-20: eps_ij = sqrt(eps_i eps_j), sigma_ij =
// (sigma_i + sigma_j)/2, c0 = 48 eps sigma^12, c1 = 24 eps sigma^6)
template <typename Rl>
struct SStoreT {
  LView L;
  const XRec<Rl> *xr;
  double *pf;
  const typename Pack4<Rl>::T *pk, *kpk;
  const int16_t *kidx;
  const int *ptype;  // the ion's atom types (global memory; BMH is ion-only)
  typedef int H;
  typedef double P;  // position and force type of the term routines
  __device__ H h(int a) const { return a; }
  // per-atom parameters: the ion + vapor range has its own, gas atoms read
  // their species atom's
  __device__ typename Pack4<Rl>::T par(H a) const { return a < L.gbase ? pk[a] : kpk[kidx[a]]; }
  __device__ void pos(H a, double p[3]) const { p[0] = xr[a].x[0]; p[1] = xr[a].x[1]; p[2] = xr[a].x[2]; }
  __device__ void fadd(H a, double fx, double fy, double fz) const {
    atomicAdd(&pf[3 * a], fx); atomicAdd(&pf[3 * a + 1], fy); atomicAdd(&pf[3 * a + 2], fz);
  }
  __device__ double w(H a) const { return (double)xr[a].w; }
  __device__ double q(H a) const { return (double)par(a).x; }
  // .y = sqrt(24 eps), .z = sigma / 2 (batch_load)
  __device__ int type(H a) const { return ptype[a]; }
  template <typename R2>
  __device__ void coeff(const DevReplica &, H a, H b, R2 &c0, R2 &c1) const {
    const typename Pack4<Rl>::T pa = par(a), pb = par(b);
    R2 e24 = (R2)pa.y * (R2)pb.y;
    R2 sg = (R2)pa.z + (R2)pb.z;
    R2 s2 = sg * sg, s6 = s2 * s2 * s2;
    c1 = e24 * s6;
    c0 = (R2)2 * c1 * s6;
  }
};

// AMBER non-excluded ion pairs (computeLong, compute_AMBER.cpp:21-49), vi
// (VaporIon.cpp:5-43) and vv (VaporVapor.cpp:5-45) as one pass over the ion
// + in-domain vapor atoms, each unordered pair evaluated once (Newton's third
// law).  The ion range [0, ni) and the vapor range [ni, nh) are each cut into
// 64-atom blocks (the last one of each range partial), and the work units are
// block pairs of three kinds: ion-ion (bi <= bj), ion-vapor (every ion block
// with every vapor block) and vapor-vapor (bi <= bj).  A wave takes a unit
// with lane l owning atom i = start_i + l and, at step k, partner
// j = start_j + (l + k) mod 64, so the 64 lanes always touch 64 different j.
// F_i accumulates in FP64 registers.  F_j accumulates in FP64 registers that
// travel with their atom: the partner of lane l at step k + 1 is the partner
// of lane l + 1 at step k, so after each step every lane takes the
// accumulator of lane l + 1 (a DPP wave rotate), and each accumulator is
// added to LDS once at the end of a work item.  A diagonal unit runs
// k = 1..32 (k = 32 on half the lanes), which visits each of its pairs once;
// an off-diagonal one runs k and k + 32 together (two independent chains per
// lane), k = 0..31.  Ion-ion pairs are selected by the AMBER exclusion
// bitmask, ion-vapor pairs by vi, vapor-vapor pairs of distinct molecules by
// vv.  A work item reads its lane's 64 selection bits for the partner block
// once (real partner atom, and for ion-ion pairs not excluded), rotated so
// that bit t belongs to the item's step t; F_i and the F_j accumulators take
// s (x_i - x_j) by FP64 fused multiply-adds, and each step's partner data is
// read from LDS one step ahead.  For an ion of a few atoms the ion-vapor
// units are instead one per vapor block: each lane holds a vapor atom and
// loops over the ion atoms, whose forces are wave sums.
//
// Work items are runs of L pairs per lane of one unit (L = 32, 16 or 8,
// whichever gives the waves the fewest rounds), dealt round-robin to the
// waves, so the waves finish together.  The LDS atom arrays hold finite
// values one block past the ion + vapor range (batch_load), so every lane
// loads and adds unconditionally and the loop body has no branches: masked
// pairs carry a zero force.

// the work item's atom i (lane)
template <typename Rl>
struct DenseI {
  double x0, x1, x2;
  Rl qi, sei, sgi, mi, wi;
  uint32_t ok;  // the lane's atom is real
};

// Unit kinds: the pair selection that cannot change within a unit is
// compiled out.
// kPairIonVapFull: ion-vapor units of two whole blocks, where every partner
// is real and none is excluded, so no selection at all.
enum { kPairIonIon = 1, kPairIonVap = 2, kPairVapVap = 3, kPairIonVapFull = 4 };

// one pair (i, j) with its selection bit from the unit's partner mask:
// returns s with the pair force on i = s (x_i - x_j) and on j the opposite,
// zero for masked pairs; the FP64 displacement x_i - x_j in (dx, dy, dz);
// energies by reference label
// a partner's LDS data
template <typename Rl>
struct DenseJ {
  typename Pack4<Rl>::T p;
  double x0, x1, x2;
  Rl w;
};

template <typename Rl>
__device__ inline DenseJ<Rl> dense_load(const BatchLds<Rl> &B, int j) {
  DenseJ<Rl> d;
  d.p = B.pk[j];
  const XRec<Rl> &xq = B.xr[j];
  d.x0 = xq.x[0]; d.x1 = xq.x[1]; d.x2 = xq.x[2];
  d.w = xq.w;
  return d;
}

template <typename Rl, bool EF, int CLS>
__device__ inline double dense_pair(const DenseJ<Rl> &q, const DenseI<Rl> &c, uint32_t pass, double &dx,
                                    double &dy, double &dz, double &e_ion, double &e_vi, double &e_vv) {
  const typename Pack4<Rl>::T pj = q.p;
  dx = c.x0 - q.x0;
  dy = c.x1 - q.x1;
  dz = c.x2 - q.x2;
  const Rl wj = q.w;
  // vapor-vapor pairs of one molecule are intramolecular (integer test, no
  // short-circuit control flow)
  const uint32_t on = CLS == kPairVapVap ? pass & (pj.w != c.mi ? 1u : 0u) : CLS == kPairIonVapFull ? 1u : pass;
  // r^2 from the FP64 displacement, rounded once to pair precision
  const Rl r2 = (Rl)(dx * dx + dy * dy + dz * dz);
  const Rl rsq = on ? r2 : (Rl)1;
  // 24 eps_ij and sigma_ij from the packed per-atom sqrt(24 eps) and
  // sigma / 2; with u3 = (sigma_ij / r)^6 the reference's r^-6 (c0 r^-6 -
  // c1) is 24 eps u3 (2 u3 - 1)
  const Rl e24 = c.sei * pj.y;
  const Rl sg = c.sgi + pj.z;
  const Rl w = CLS == kPairIonIon ? (Rl)1 : c.wi * wj;
  const Rl wm = on ? w : (Rl)0;
  Rl r2inv, rinv;
  inv_r<Rl>(rsq, r2inv, rinv);
  const Rl u = sg * sg * r2inv, u3 = u * u * u;
  const Rl flj = e24 * u3 * ((Rl)2 * u3 - (Rl)1);
  const Rl fc = c.qi * pj.x * rinv;
  if (EF) {
    const double ep = on ? (double)(e24 * (Rl)(1.0 / 6.0) * u3 * (u3 - (Rl)1) + fc) : 0.0;
    (CLS == kPairIonIon ? e_ion : CLS == kPairVapVap ? e_vv : e_vi) += ep;
  }
  // the pair-precision scalar force over r, applied to the FP64
  // displacement by the caller
  return (double)((flj + fc) * r2inv * wm);
}

// one bit per partner slot, rotated by r at the start of a work item so that
// bit t belongs to the item's step t
__device__ inline uint64_t rotr64(uint64_t v, int r) {
  r &= 63;
  return r ? (v >> r) | (v << (64 - r)) : v;
}

// the off-diagonal run of one work item: partners k and k + co together
// (two independent chains per lane), each with its travelling F_j
// accumulator; returns through the references.  Partner slot of lane l at
// step k: (l + k) & jm.  A full partner block (jm = 63, co = 32) takes 64
// steps; a partner block whose atoms all sit in its first 32 slots takes
// 32 (jm = 31, co = 16): lanes l and l + 32 then meet the same partners, and
// the wave rotate still hands each accumulator to the lane that meets its
// partner next, because lane 31 receives lane 32's and lane 63 lane 0's.
// Selection bits: bit t of sa for the partner of step kb + t, of sb for that
// of step kb + t + co (a run has at most 16 steps).
template <typename Rl, bool EF, int CLS>
__device__ inline void dense_run_off(const BatchLds<Rl> &B, const DenseI<Rl> &c, int lane, int jb, int kb, int hL,
                                     int jm, int co, uint32_t sa, uint32_t sb, double &fx, double &fy, double &fz,
                                     double *a0, double *a1, double &e_ion, double &e_vi, double &e_vv) {
  auto step = [&](const DenseJ<Rl> &qa, const DenseJ<Rl> &qb, int t) {
    const uint32_t pa = (sa >> t) & 1u, pb = (sb >> t) & 1u;
    double dx, dy, dz, ex, ey, ez;
    const double s = dense_pair<Rl, EF, CLS>(qa, c, pa, dx, dy, dz, e_ion, e_vi, e_vv);
    const double u = dense_pair<Rl, EF, CLS>(qb, c, pb, ex, ey, ez, e_ion, e_vi, e_vv);
    fx = fma(s, dx, fx); fy = fma(s, dy, fy); fz = fma(s, dz, fz);
    fx = fma(u, ex, fx); fy = fma(u, ey, fy); fz = fma(u, ez, fz);
    a0[0] = wave_rol1(fma(-s, dx, a0[0]));
    a0[1] = wave_rol1(fma(-s, dy, a0[1]));
    a0[2] = wave_rol1(fma(-s, dz, a0[2]));
    a1[0] = wave_rol1(fma(-u, ex, a1[0]));
    a1[1] = wave_rol1(fma(-u, ey, a1[1]));
    a1[2] = wave_rol1(fma(-u, ez, a1[2]));
  };
  // partners of step t: slots (lane + kb + t) & jm and (lane + kb + t + co) & jm;
  // the loads of the next step are issued before this step's arithmetic
  // (the last prefetch wraps within the block: in bounds, unused)
  const int l0 = lane + kb;
  DenseJ<Rl> qa = dense_load(B, jb + (l0 & jm)), qb = dense_load(B, jb + ((l0 + co) & jm));
  for (int t = 0; t < hL; t += 2) {
    const DenseJ<Rl> ra = dense_load(B, jb + ((l0 + t + 1) & jm)), rb = dense_load(B, jb + ((l0 + t + 1 + co) & jm));
    // keep the loads ahead of the arithmetic that does not need them
    __builtin_amdgcn_sched_barrier(0);
    step(qa, qb, t);
    qa = dense_load(B, jb + ((l0 + t + 2) & jm));
    qb = dense_load(B, jb + ((l0 + t + 2 + co) & jm));
    __builtin_amdgcn_sched_barrier(0);
    step(ra, rb, t + 1);
  }
}

// Item filter: every block pair, only those of two whole-ion blocks (ion-ion
// pairs: no vapor position or weight involved), or all but those.
enum { kDenseAll = 0, kDenseIonOnly = 1, kDenseNotIonOnly = 2 };

// ions of at most this many atoms take the one-unit-per-vapor-block form
constexpr int kIonBroadcastMax = 16;

// returns the number of work items (dealt round-robin from wave 0)
template <typename Rl, bool EF>
__device__ inline int dense_n3(const DevReplica &R, const BatchLds<Rl> &B, int nh, int ni, bool am, bool vi, bool vv,
                               int t, int t0, int nt, int filt, double *e) {
  const Rl qq = (Rl)R.qqrd2e;
  // threads [t0, t0 + nt) take part (whole waves).  The wave index and the
  // counts are uniform within a wave: say so, so that the item loops and
  // their trip counts stay scalar
  const int lane = t & 63, wave = __builtin_amdgcn_readfirstlane((t - t0) >> 6);
  const int nw = __builtin_amdgcn_readfirstlane(nt >> 6);
  nh = __builtin_amdgcn_readfirstlane(nh);
  ni = __builtin_amdgcn_readfirstlane(ni);
  const int NBI = (ni + 63) >> 6, NBV = (nh - ni + 63) >> 6;
  const bool doII = am && ni >= 2 && filt != kDenseNotIonOnly;
  const bool doIV = vi && ni > 0 && NBV > 0 && filt != kDenseIonOnly;
  const bool doVV = vv && NBV > 0 && filt != kDenseIonOnly;
  const bool bcast = ni <= kIonBroadcastMax;
  // a range's last block may be partial; one with at most 32 atoms is a
  // half-width partner block (32 steps instead of 64)
  const bool ihalf = (ni & 63) != 0 && (ni & 63) <= 32;
  const int nvap = nh - ni;
  const bool vhalf = (nvap & 63) != 0 && (nvap & 63) <= 32;
  // off-diagonal units: full (64 steps) and half (32 steps)
  int n_full = 0, n_half = 0;
  if (doII) {
    n_full += (NBI - 1) * (NBI - 2) / 2 + (ihalf ? 0 : NBI - 1);
    n_half += ihalf ? NBI - 1 : 0;
  }
  if (doIV && !bcast)
    for (int bi = 0; bi < NBI; bi++)
      for (int bj = 0; bj < NBV; bj++)
        ((bj == NBV - 1 && vhalf) || (bi == NBI - 1 && ihalf) ? n_half : n_full)++;
  if (doVV) {
    n_full += (NBV - 1) * (NBV - 2) / 2 + (vhalf ? 0 : NBV - 1);
    n_half += vhalf ? NBV - 1 : 0;
  }
  const int n_diag = (doII ? NBI : 0) + (doVV ? NBV : 0) + n_half;
  const int n_off = n_full;
  const int n_bc = doIV && bcast ? NBV : 0;
  // run length: the fewest rounds of items per wave, counting about six
  // steps of set-up and flush per item
  int L = 32, best = 1 << 30;
  for (int l = 32; l >= 8; l >>= 1) {
    const int items = n_diag * (32 / l) + n_off * (64 / l) + n_bc, c = (items + nw - 1) / nw * (l + 6);
    if (c < best) { best = c; L = l; }
  }
  const int W = __builtin_amdgcn_readfirstlane(R.amber_words);
  double *__restrict__ pf = B.f;
  double e_ion = 0, e_vi = 0, e_vv = 0;
  // the lane's atom of a unit whose i block starts at ib
  auto setup = [&](int i, int ilim, DenseI<Rl> &c) {
    c.x0 = B.xr[i].x[0]; c.x1 = B.xr[i].x[1]; c.x2 = B.xr[i].x[2];
    const typename Pack4<Rl>::T pi = B.pk[i];
    c.qi = qq * pi.x; c.sei = pi.y; c.sgi = pi.z; c.mi = pi.w; c.wi = B.xr[i].w;
    c.ok = i < ilim ? 1u : 0u;
  };
  // selection bits of the lane's partners in the j block at jb, bit s for
  // slot s: a real atom (jb + s < jlim), for ion-ion pairs not excluded
  // (row i of the AMBER bitmask), none for a padding lane
  auto partner_bits = [&](int cls, int i, const DenseI<Rl> &c, int jb, int jlim) -> uint64_t {
    const int jv = min(max(jlim - jb, 0), 64);
    uint64_t V = jv >= 64 ? ~0ull : ((1ull << jv) - 1ull);
    if (cls == kPairIonIon) {
      const uint32_t *row = B.allow + (size_t)(i < ni ? i : 0) * W;
      const int w0 = jb >> 5, wm = W > 0 ? W - 1 : 0;
      V &= (uint64_t)row[min(w0, wm)] | ((uint64_t)row[min(w0 + 1, wm)] << 32);
    }
    return c.ok ? V : 0ull;
  };
  auto flush = [&](int idx, double ax, double ay, double az) {
    atomicAdd(&pf[3 * idx], ax);
    atomicAdd(&pf[3 * idx + 1], ay);
    atomicAdd(&pf[3 * idx + 2], az);
  };
  // one block-pair unit's run: i block at ib (atoms < ilim real), j block at
  // jb (atoms < jlim real)
  auto run = [&](int cls, bool diag, bool half, int ib, int ilim, int jb, int jlim, int part) {
    const int i = ib + lane;
    DenseI<Rl> c;
    setup(i, ilim, c);
    uint64_t V = partner_bits(cls, i, c, jb, jlim);
    double fx = 0, fy = 0, fz = 0;
    if (diag) {
      // k = 1..32 in runs of L; k = 32 only on lanes 0..31 (lane l >= 32
      // meets slot l - 32 at k = 32, which lane l - 32 visits at k = 32)
      const int kb = 1 + part * L;
      if (lane >= 32) V &= ~(1ull << (lane - 32));
      const uint32_t sa = (uint32_t)rotr64(V, (lane + kb) & 63);  // bit t: step kb + t (L <= 32)
      double a0x = 0, a0y = 0, a0z = 0;
      auto step = [&](const DenseJ<Rl> &q, int t) {
        const uint32_t pa = (sa >> t) & 1u;
        double dx, dy, dz, sv;
        if (cls == kPairIonIon) sv = dense_pair<Rl, EF, kPairIonIon>(q, c, pa, dx, dy, dz, e_ion, e_vi, e_vv);
        else sv = dense_pair<Rl, EF, kPairVapVap>(q, c, pa, dx, dy, dz, e_ion, e_vi, e_vv);
        fx = fma(sv, dx, fx); fy = fma(sv, dy, fy); fz = fma(sv, dz, fz);
        a0x = wave_rol1(fma(-sv, dx, a0x));
        a0y = wave_rol1(fma(-sv, dy, a0y));
        a0z = wave_rol1(fma(-sv, dz, a0z));
      };
      // one step of prefetch, as in the off-diagonal runs (L is even)
      const int l0 = lane + kb;
      DenseJ<Rl> q0 = dense_load(B, jb + (l0 & 63));
      for (int t = 0; t < L; t += 2) {
        const DenseJ<Rl> q1 = dense_load(B, jb + ((l0 + t + 1) & 63));
        __builtin_amdgcn_sched_barrier(0);
        step(q0, t);
        q0 = dense_load(B, jb + ((l0 + t + 2) & 63));
        __builtin_amdgcn_sched_barrier(0);
        step(q1, t + 1);
      }
      // each accumulator now belongs to the partner one step past the run
      // (64 distinct atoms of the j block; zero sums for masked partners)
      flush(jb + ((lane + kb + L) & 63), a0x, a0y, a0z);
    } else {
      // two independent partners per step, k and k + co, each with its own
      // travelling F_j accumulator
      const int hL = L >> 1, kb = part * hL;
      const int jm = half ? 31 : 63, co = half ? 16 : 32;
      // a half block's 32 slots repeat in both halves, so a rotate by r < 32
      // rotates each half; bits co.. then hold the partners k + co
      if (half) V = (V & 0xffffffffull) | (V << 32);
      V = rotr64(V, (lane + kb) & jm);
      const uint32_t sa = (uint32_t)V, sb = (uint32_t)(V >> co);
      double a0[3] = {0, 0, 0}, a1[3] = {0, 0, 0};
      if (cls == kPairIonIon)
        dense_run_off<Rl, EF, kPairIonIon>(B, c, lane, jb, kb, hL, jm, co, sa, sb, fx, fy, fz, a0, a1, e_ion, e_vi,
                                           e_vv);
      else if (cls == kPairIonVapFull)
        dense_run_off<Rl, EF, kPairIonVapFull>(B, c, lane, jb, kb, hL, jm, co, sa, sb, fx, fy, fz, a0, a1, e_ion,
                                               e_vi, e_vv);
      else if (cls == kPairIonVap)
        dense_run_off<Rl, EF, kPairIonVap>(B, c, lane, jb, kb, hL, jm, co, sa, sb, fx, fy, fz, a0, a1, e_ion, e_vi,
                                           e_vv);
      else
        dense_run_off<Rl, EF, kPairVapVap>(B, c, lane, jb, kb, hL, jm, co, sa, sb, fx, fy, fz, a0, a1, e_ion, e_vi,
                                           e_vv);
      flush(jb + ((lane + kb + hL) & jm), a0[0], a0[1], a0[2]);
      flush(jb + ((lane + kb + hL + co) & jm), a1[0], a1[1], a1[2]);
    }
    if (i < ilim) flush(i, fx, fy, fz);
  };
  int item = 0;
  // ion-ion units (the partial last block, if any, on the partner side)
  if (doII)
    for (int bi = 0; bi < NBI; bi++)
      for (int bj = bi; bj < NBI; bj++) {
        const bool half = bi != bj && bj == NBI - 1 && ihalf;
        const int np = (bi == bj || half ? 32 : 64) / L;
        for (int part = 0; part < np; part++, item++)
          if (item % nw == wave) run(kPairIonIon, bi == bj, half, 64 * bi, ni, 64 * bj, ni, part);
      }
  // ion-vapor units; a half-width block goes to the partner side (the ion
  // block when only it is one: the pair form is symmetric in i and j)
  if (doIV && !bcast)
    for (int bi = 0; bi < NBI; bi++)
      for (int bj = 0; bj < NBV; bj++) {
        const bool vh = bj == NBV - 1 && vhalf, ih = bi == NBI - 1 && ihalf, swap = ih && !vh;
        const int np = (vh || ih ? 32 : 64) / L;
        for (int part = 0; part < np; part++, item++)
          if (item % nw == wave) {
            if (swap) run(kPairIonVap, false, true, ni + 64 * bj, nh, 64 * bi, ni, part);
            else if (!vh && 64 * bi + 64 <= ni && ni + 64 * bj + 64 <= nh)
              run(kPairIonVapFull, false, false, 64 * bi, ni, ni + 64 * bj, nh, part);
            else run(kPairIonVap, false, vh, 64 * bi, ni, ni + 64 * bj, nh, part);
          }
      }
  if (doIV && bcast)
    for (int bj = 0; bj < NBV; bj++, item++) {
      if (item % nw != wave) continue;
      // the lane's vapor atom against every ion atom (uniform partner):
      // F_vapor in registers, F_ion a wave sum per ion atom
      const int v = ni + 64 * bj + lane;
      DenseI<Rl> c;
      setup(v, nh, c);
      double fx = 0, fy = 0, fz = 0;
      for (int i = 0; i < ni; i++) {
        double dx, dy, dz;
        const double sv = dense_pair<Rl, EF, kPairIonVap>(dense_load(B, i), c, c.ok, dx, dy, dz, e_ion, e_vi, e_vv);
        fx = fma(sv, dx, fx); fy = fma(sv, dy, fy); fz = fma(sv, dz, fz);
        const double sx = wave_sum_full(-sv * dx), sy = wave_sum_full(-sv * dy), sz = wave_sum_full(-sv * dz);
        if (lane == 0) flush(i, sx, sy, sz);
      }
      if (v < nh) flush(v, fx, fy, fz);
    }
  // vapor-vapor units
  if (doVV)
    for (int bi = 0; bi < NBV; bi++)
      for (int bj = bi; bj < NBV; bj++) {
        const bool half = bi != bj && bj == NBV - 1 && vhalf;
        const int np = (bi == bj || half ? 32 : 64) / L;
        for (int part = 0; part < np; part++, item++)
          if (item % nw == wave) run(kPairVapVap, bi == bj, half, ni + 64 * bi, nh, ni + 64 * bj, nh, part);
      }
  if (EF) { e[U_ION] += e_ion; e[U_VI] += e_vi; e[U_VV] += e_vv; }
  return item;
}

template <typename Rl>
__device__ int dense_heavy(const DevReplica &R, const BatchLds<Rl> &B, const LView &L, int t, int t0, int nt,
                           int filt, bool eflag, double *e) {
  const int nh = L.gbase, ni = L.n_ion;
  const bool am = (R.flags & B_AMBER) != 0, vi = (R.flags & B_VI) != 0, vv = (R.flags & B_VV) != 0;
  if (!(am || vi || vv) || nh == 0) return 0;
  if (eflag) return dense_n3<Rl, true>(R, B, nh, ni, am, vi, vv, t, t0, nt, filt, e);
  return dense_n3<Rl, false>(R, B, nh, ni, am, vi, vv, t, t0, nt, filt, e);
}

// the batched path's force phase: registration order of setPotential.cpp
// with vi, vv and the AMBER non-bonded pairs in the dense pass
// MB: the kernel carries the many-body and BMH ion terms (Tersoff, SW,
// Born-Mayer-Huggins).  Without them the compiler keeps every value in
// registers (no spills); a launch takes the MB variant only when one of its
// replicas selects such a term.
template <typename Rl, bool MB>
__device__ void batch_terms(const DevReplica &R, const SStoreT<Rl> &s, const BatchLds<Rl> &B, int t, int nt,
                            bool eflag, double *e) {
  const uint32_t bits = R.flags;
  if (bits & B_GI) term_gi<Rl>(R, s, t, nt, eflag, e);
  dense_heavy<Rl>(R, B, s.L, t, 0, nt, kDenseAll, eflag, e);
  if (bits & B_GG) term_gg<Rl>(R, s, t, nt, eflag, e);
  if (bits & B_GV) term_gv<Rl>(R, s, t, nt, eflag, e);
  if (bits & B_EFIELD) term_efield(R, s, t, nt);
  if (MB && (bits & B_TERSOFF)) term_tersoff(R, s, t, nt, eflag, e);
  if (MB && (bits & B_SW)) term_sw(R, s, t, nt, eflag, e);
  if (bits & B_AMBER) term_amber_bonded(R, s, t, nt, eflag, e);
  if (MB && (bits & B_BMH)) term_born(R, s, t, nt, eflag, e);
  if (bits & B_VINTRA) term_vintra(R, s, t, nt, eflag, e);
  if (bits & B_GINTRA) term_gintra(R, s, t, nt, eflag, e);
}

// The forces of a step whose vapor SHAKE runs beside them: phase A uses
// only ion and gas positions (threads t of [t0, t0 + nt), whole waves);
// phase B (all threads, after SHAKE and the vapor weights) the rest.
template <typename Rl, bool MB>
__device__ void batch_terms_a(const DevReplica &R, const SStoreT<Rl> &s, const BatchLds<Rl> &B, int t, int t0, int nt,
                              bool eflag, double *e) {
  const uint32_t bits = R.flags;
  const int tl = t - t0;
  if (bits & B_GI) term_gi<Rl>(R, s, tl, nt, eflag, e);
  dense_heavy<Rl>(R, B, s.L, t, t0, nt, kDenseIonOnly, eflag, e);
  if (bits & B_EFIELD) term_efield(R, s, tl, nt);
  if (MB && (bits & B_TERSOFF)) term_tersoff(R, s, tl, nt, eflag, e);
  if (MB && (bits & B_SW)) term_sw(R, s, tl, nt, eflag, e);
  if (bits & B_AMBER) term_amber_bonded(R, s, tl, nt, eflag, e);
  if (MB && (bits & B_BMH)) term_born(R, s, tl, nt, eflag, e);
  if (bits & B_GINTRA) term_gintra(R, s, tl, nt, eflag, e);
}
template <typename Rl>
__device__ void batch_terms_b(const DevReplica &R, const SStoreT<Rl> &s, const BatchLds<Rl> &B, int t, int nt,
                              bool eflag, double *e) {
  const uint32_t bits = R.flags;
  const int items = dense_heavy<Rl>(R, B, s.L, t, 0, nt, kDenseNotIonOnly, eflag, e);
  // the remaining terms start on the first wave that got one item fewer
  const int tr = (t + nt - 64 * (items % (nt >> 6))) % nt;
  if (bits & B_GG) term_gg<Rl>(R, s, tr, nt, eflag, e);
  if (bits & B_GV) term_gv<Rl>(R, s, tr, nt, eflag, e);
  if (bits & B_VINTRA) term_vintra(R, s, tr, nt, eflag, e);
}

template <typename Rl>
__device__ inline void batch_load(const DevReplica &R, const LView &L, const BatchLds<Rl> &B) {
  const double C = 0.5 * R.dt * kKickUnit;
  if (R.flags & B_AMBER)
    for (int k = threadIdx.x; k < L.n_ion * R.amber_words; k += blockDim.x) B.allow[k] = R.amber_allow[k];
  if (threadIdx.x == 0) shake_lds_fill(R, R.sp[0], B.shk, 0);
  if (threadIdx.x == blockDim.x - 1) shake_lds_fill(R, R.sp[2], B.shk, 1);
  const int nkt = L.n_ion + L.na_v + L.na_g;
  for (int k = threadIdx.x; k < nkt; k += blockDim.x) {
    int sp = k < L.n_ion ? 0 : (k < L.n_ion + L.na_v ? 2 : 1);
    int idx = sp == 0 ? k : (sp == 2 ? k - L.n_ion : k - L.n_ion - L.na_v);
    const DevSpecies &S = R.sp[sp];
    B.kct[k] = C / S.mass[idx];
    B.kms[k] = S.mass[idx];
    B.kpk[k].x = (Rl)S.charge[idx];
    B.kpk[k].y = (Rl)(S.se[idx] * kSqrt24);
    B.kpk[k].z = (Rl)(0.5 * S.sg[idx]);
    B.kpk[k].w = (Rl)-1;
  }
  for (int a = threadIdx.x; a < L.n_in; a += blockDim.x) {
    int sp, idx, kk;
    if (a < L.n_ion) { sp = 0; idx = a; kk = a; }
    else if (a < L.gbase) { sp = 2; idx = (a - L.vbase) % L.na_v; kk = L.n_ion + idx; }
    else { sp = 1; idx = (a - L.gbase) % L.na_g; kk = L.n_ion + L.na_v + idx; }
    size_t g = (size_t)gatom(R, L, a);
    for (int q = 0; q < 3; q++) {
      B.xr[a].x[q] = R.atom_x[3 * g + q];
      B.v[3 * a + q] = R.atom_v[3 * g + q];
      B.f[3 * a + q] = R.atom_f[3 * g + q];
    }
    const DevSpecies &S = R.sp[sp];
    B.kidx[a] = (int16_t)kk;
    if (a < L.gbase) {
      const int m = sp == 2 ? (a - L.vbase) / L.na_v : -1;
      B.pk[a].x = (Rl)S.charge[idx];
      B.pk[a].y = (Rl)(S.se[idx] * kSqrt24);
      B.pk[a].z = (Rl)(0.5 * S.sg[idx]);
      B.pk[a].w = (Rl)m;
    }
    B.xr[a].w = (Rl)R.atom_w[g];
  }
  // finite values for one block past the ion + vapor range (read by
  // dense_n3): parameters past the range, atom data past n_in
  const int hend = L.gbase + 64;
  for (int a = L.gbase + threadIdx.x; a < hend; a += blockDim.x) {
    B.pk[a].x = (Rl)0; B.pk[a].y = (Rl)0; B.pk[a].z = (Rl)0; B.pk[a].w = (Rl)-1;
  }
  for (int a = L.n_in + threadIdx.x; a < hend; a += blockDim.x) {
    for (int q = 0; q < 3; q++) { B.xr[a].x[q] = 0.0; B.f[3 * a + q] = 0.0; }
    B.xr[a].w = (Rl)0;
  }
}

template <typename Rl>
__device__ inline void batch_store(const DevReplica &R, const LView &L, const BatchLds<Rl> &B) {
  for (int a = threadIdx.x; a < L.n_in; a += blockDim.x) {
    size_t g = (size_t)gatom(R, L, a);
    for (int q = 0; q < 3; q++) {
      R.atom_x[3 * g + q] = B.xr[a].x[q];
      R.atom_v[3 * g + q] = B.v[3 * a + q];
      R.atom_f[3 * g + q] = B.f[3 * a + q];
    }
    R.atom_w[g] = (double)B.xr[a].w;
  }
}

template <typename Rl, bool MB>
__global__ void __launch_bounds__(512) k_batched(const DevReplica *__restrict__ reps,
                                                 const int *__restrict__ order, int n_rep, int nsteps,
                                                 int eflag_every, long long step0, int cap, int cap_h, int kcap,
                                                 int acap, int scap, int4 *__restrict__ cnt) {
  extern __shared__ __align__(16) char smem[];
  BatchLds<Rl> B = carve<Rl>(smem, cap, cap_h, kcap, acap, scap);
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int rr = blockIdx.x; rr < n_rep; rr += gridDim.x) {
    const DevReplica &R = reps[order[rr]];
    Dyn *D = R.dyn;
    {
      if (D->status & ST_LDS_CAP) goto report;
      LView L = view_from_dyn(R);
      if (L.n_in > cap || L.gbase + 64 > cap_h || L.n_ion + L.na_v + L.na_g > kcap ||
          ((R.flags & B_AMBER) && L.n_ion * R.amber_words > acap)) {
        if (tid == 0) atomicOr(&D->status, ST_LDS_CAP);
        goto report;
      }
      batch_load(R, L, B);
      int lp = D->loopPair, lu = D->loop_update;
      long long acc_steps = 0;
      const DevSpecies &I = R.sp[0];
      const bool shake = R.shake != 0;
      // ion thermostat (nvt.cpp): the mode is uniform over the launch, zeta
      // lives in a register and reaches the replica with the state
      const int thermo = D->thermo;
      const double Tnh = D->Tnh_ion;
      double zeta = D->zeta_ion;
      const long long step_base = D->step;
      // SHAKE diagnostics accumulate in LDS (sint[32]: max iterations,
      // sint[33]: failures) and reach the replica's counters once per launch
      if (tid == 0) { B.sint[32] = 0; B.sint[33] = 0; }
      __syncthreads();
      bool dead = false;
      // under NVE the second half-kick of a step is carried into the first of
      // the next (the same per-atom operations in the same order: nothing
      // between them reads the velocities); the last step of a launch applies
      // its own
      bool pend = false;
      for (int s = 0; s < nsteps && !dead; s++) {
        const bool eflag = eflag_every > 0 && ((step0 + s + 1) % eflag_every) == 0;
        // MD::velocity_scaling (verlet.cpp:22): the ion is kicked first, its
        // temperature reduced and its velocities scaled to 300 K; the fused
        // loop below then only drifts it (kick coefficient 0)
        const bool vscale = thermo == THERMO_SCALE;
        if (vscale) {
          double k2 = 0;
          for (int a = tid; a < L.n_ion; a += nt) {
            double kc = B.kct[B.kidx[a]], m = B.kms[a];
            for (int q = 0; q < 3; q++) {
              double v = B.v[3 * a + q] + B.f[3 * a + q] * kc;
              B.v[3 * a + q] = v;
              k2 += v * v * m;
            }
          }
          k2 = wave_sum(k2);
          if ((tid & 63) == 0) B.red[tid >> 6] = k2;
          __syncthreads();
          double sv = 0;
          for (int w = 0; w < nt / 64; w++) sv += B.red[w];
          const double ratio = sqrt(300.0 / ion_temperature(sv, I.natoms));
          for (int a = tid; a < L.n_ion; a += nt)
            for (int q = 0; q < 3; q++) B.v[3 * a + q] *= ratio;
        }
        // first half-kick + drift + force clear (verlet.cpp:21-25), after the
        // previous step's second half-kick when it was carried over
        for (int a = tid; a < L.n_in; a += nt) {
          const double kf = B.kct[B.kidx[a]];
          double kc = vscale && a < L.n_ion ? 0.0 : kf;
          bool keep = shake && a < L.gbase;
          for (int q = 0; q < 3; q++) {
            double v = B.v[3 * a + q];
            if (pend) v = v + B.f[3 * a + q] * kf;
            v = v + B.f[3 * a + q] * kc;
            B.v[3 * a + q] = v;
            if (keep) B.xo[3 * a + q] = B.xr[a].x[q];
            B.xr[a].x[q] += v * R.dt;
            B.f[3 * a + q] = 0.0;
          }
        }
        __syncthreads();
        SStoreT<Rl> ss;
        ss.L = L; ss.xr = B.xr; ss.pf = B.f; ss.pk = B.pk; ss.kpk = B.kpk; ss.kidx = B.kidx;
        ss.ptype = R.sp[0].type;
        double e[U_N] = {0, 0, 0, 0, 0, 0, 0, 0};
        // SHAKE of one molecule whose first local atom is base
        auto shake_one = [&](const DevSpecies &S, int base) {
          XRec<Rl> *x = B.xr + base;
          double *v = B.v + 3 * base;
          const double *xo = B.xo + 3 * base;
          auto X = [&](int i, int q) -> double & { return x[i].x[q]; };
          auto XO = [&](int i, int q) -> double { return xo[3 * i + q]; };
          auto V = [&](int i, int q) -> double & { return v[3 * i + q]; };
          const int sidx = &S == &R.sp[0] ? 0 : 1;
          int it = shake_molecule_lds(R, B.shk, sidx, X, XO, R.shake_max_iter);
          if (it < 0) {
            atomicAdd(&B.sint[33], 1);
            atomicMax(&B.sint[32], -it);
            shake_undo(S, X, XO, V, R.dt);
            return;
          }
          atomicMax(&B.sint[32], it);
          if (R.shake == 1) shake_velocity_lds(B.shk, sidx, S.natoms, X, XO, V, R.dt);
        };
        // the same for constraint group g of G of a vapor molecule (one lane
        // per group): the group's own SHAKE velocities; the first lane counts
        // and, on a failure, undoes the whole molecule
        auto shake_grp = [&](const DevSpecies &S, int base, int g, int G) {
          XRec<Rl> *x = B.xr + base;
          double *v = B.v + 3 * base;
          const double *xo = B.xo + 3 * base;
          auto X = [&](int i, int q) -> double & { return x[i].x[q]; };
          auto XO = [&](int i, int q) -> double { return xo[3 * i + q]; };
          auto V = [&](int i, int q) -> double & { return v[3 * i + q]; };
          int it = shake_group_lds(R, B.shk, 1, g, G, X, XO, R.shake_max_iter);
          if (it < 0) {
            if (g == 0) {
              atomicAdd(&B.sint[33], 1);
              atomicMax(&B.sint[32], -it);
              shake_undo(S, X, XO, V, R.dt);
            }
            return;
          }
          if (g == 0) atomicMax(&B.sint[32], it);
          if (R.shake == 1) {
            const double inv_dt = 1.0 / R.dt;
            const unsigned long long gm = B.shk.gm[4 + g];
            for (int a = 0; a < S.natoms; a++)
              if ((gm >> a) & 1ull)
                for (int q = 0; q < 3; q++) V(a, q) = (X(a, q) - XO(a, q)) * inv_dt;
          }
        };
        const DevSpecies &Vs = R.sp[2];
        const int nwv = nt >> 6;
        // A step with constrained vapor, an unconstrained ion and no rebuild
        // due: the vapor SHAKE (one thread per molecule, latency-bound) runs
        // on the last wave(s) while the other waves compute the ion COM, the
        // gas weights and the forces that involve no vapor position (phase A);
        // then the vapor weights and the remaining forces (phase B).  The
        // forces are the same terms as in the sequential order; only their
        // accumulation order differs.
        const bool overlap =
            shake && Vs.nrigid && !I.nrigid && L.nv > 0 && nwv >= 2 && lp + 1 <= lu && thermo == THERMO_NVE;
        if (overlap) {
          // a vapor molecule whose constraints fall into G groups that share
          // no atom takes G adjacent lanes, one group each
          const int G = __builtin_amdgcn_readfirstlane(B.shk.ng[1]);
          const int gsh = G >= 4 ? 2 : G >= 2 ? 1 : 0;
          const int nS = min(nwv - 1, ((L.nv << gsh) + 63) >> 6), nA = nwv - nS;
          const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
          if (wave >= nA) {
            // the SHAKE chain is a long run of dependent instructions: let it
            // win instruction issue on its SIMD over the waves of force work
            __builtin_amdgcn_s_setprio(2);
            if (gsh == 0)
              for (int m = tid - 64 * nA; m < L.nv; m += 64 * nS) shake_one(Vs, L.vbase + m * L.na_v);
            else
              for (int slot = tid - 64 * nA; slot < (L.nv << gsh); slot += 64 * nS)
                shake_grp(Vs, L.vbase + (slot >> gsh) * L.na_v, slot & ((1 << gsh) - 1), 1 << gsh);
            __builtin_amdgcn_s_setprio(0);
          } else {
            // ion COM (verlet.cpp:90-107) on wave 0; no phase-A term reads a
            // weight, so the blend weights wait for the barrier
            if (wave == 0) {
              double c[6] = {0, 0, 0, 0, 0, 0};
              for (int i = tid; i < I.natoms; i += 64) {
                double m = B.kms[i];
                for (int q = 0; q < 3; q++) { c[q] += B.xr[i].x[q] * m; c[3 + q] += B.v[3 * i + q] * m; }
              }
              for (int k = 0; k < 6; k++) c[k] = wave_sum_full(c[k]) * (1 / I.molmass);
              if (tid < 6) B.sdbl[tid] = c[tid];
            }
            batch_terms_a<Rl, MB>(R, ss, B, tid, 0, 64 * nA, eflag, e);
          }
          __syncthreads();
          {
            // blend weights of the vapor and gas atoms (verlet.cpp:109-134)
            const double cx = B.sdbl[0], cy = B.sdbl[1], cz = B.sdbl[2];
            for (int a = L.n_ion + tid; a < L.n_in; a += nt) {
              double dx = B.xr[a].x[0] - cx, dy = B.xr[a].x[1] - cy, dz = B.xr[a].x[2] - cz;
              B.xr[a].w = blend_weight_r<Rl>(dx * dx + dy * dy + dz * dz);
            }
          }
          lp++;  // check_pairlist: no rebuild due (pairlist.cpp:58-63)
          __syncthreads();
          batch_terms_b<Rl>(R, ss, B, tid, nt, eflag, e);
        } else {
          if (shake) {
            // constrained molecules: the vapor and the ion (gas has no rigid
            // pairs)
            if (Vs.nrigid)
              for (int m = tid; m < L.nv; m += nt) shake_one(Vs, L.vbase + m * L.na_v);
            if (I.nrigid && tid == nt - 1) shake_one(I, 0);
            __syncthreads();
          }
          // ion COM (verlet.cpp:90-107); under Nose-Hoover also sum m v^2 of
          // the ion for MD::nosehoover_zeta (verlet.cpp:27)
          {
            const int nk = thermo == THERMO_NH ? 7 : 6;
            double c[7] = {0, 0, 0, 0, 0, 0, 0};
            for (int i = tid; i < I.natoms; i += nt) {
              double m = B.kms[i];
              for (int q = 0; q < 3; q++) { c[q] += B.xr[i].x[q] * m; c[3 + q] += B.v[3 * i + q] * m; }
              c[6] += B.v[3 * i] * B.v[3 * i] * m + B.v[3 * i + 1] * B.v[3 * i + 1] * m +
                      B.v[3 * i + 2] * B.v[3 * i + 2] * m;
            }
            for (int k = 0; k < nk; k++) {
              double sv = wave_sum(c[k]);
              if ((tid & 63) == 0) B.red[k * 16 + (tid >> 6)] = sv;
            }
            __syncthreads();
            if (tid < nk) {
              double sv = 0;
              for (int w = 0; w < nt / 64; w++) sv += B.red[tid * 16 + w];
              B.sdbl[tid] = tid < 6 ? sv * (1 / I.molmass) : sv;
            }
            __syncthreads();
            if (thermo == THERMO_NH)
              zeta = nh_zeta_step(zeta, ion_temperature(B.sdbl[6], I.natoms), Tnh, I.natoms, R.dt);
          }
          const double cx = B.sdbl[0], cy = B.sdbl[1], cz = B.sdbl[2];
          // blend weights (verlet.cpp:109-134)
          for (int a = L.n_ion + tid; a < L.n_in; a += nt) {
            double dx = B.xr[a].x[0] - cx, dy = B.xr[a].x[1] - cy, dz = B.xr[a].x[2] - cz;
            B.xr[a].w = blend_weight_r<Rl>(dx * dx + dy * dy + dz * dz);
          }
          // check_pairlist (pairlist.cpp:58-63)
          lp++;
          if (lp > lu) {
            __syncthreads();
            batch_store(R, L, B);
            if (tid < 3) { R.mol_x[tid] = B.sdbl[tid]; R.mol_v[tid] = B.sdbl[3 + tid]; }
            if (tid == 0) { D->loopPair = lp; D->zeta_ion = zeta; D->step = step_base + s; }
            __syncthreads();
            // scratch for the list tests: the positions, velocities and forces
            // (stored above, loaded again below)
            rebuild_wg(R, B.sint, B.red, reinterpret_cast<double *>(B.xr), 10 * cap);
            L = view_from_dyn(R);
            lp = D->loopPair;
            lu = D->loop_update;
            if (L.n_in > cap || L.gbase + 64 > cap_h) {
              // the stored state is step s in progress (after the rebuild):
              // the per-call path finishes it (mdn_verlet)
              if (tid == 0) { atomicOr(&D->status, ST_LDS_CAP); D->pending = 1; }
              dead = true;
              break;
            }
            batch_load(R, L, B);
            ss.L = L;
          }
          __syncthreads();
          // forces of every registered potential (verlet.cpp:31-32)
          batch_terms<Rl, MB>(R, ss, B, tid, nt, eflag, e);
        }
        __syncthreads();
        // second half-kick (verlet.cpp:34), then MD::nosehoover_ion on the ion
        // (verlet.cpp:36)
        pend = thermo == THERMO_NVE && s + 1 < nsteps;
        if (!pend) {
          const double fac = thermo == THERMO_NH ? nh_factor(zeta, R.dt) : 1.0;
          for (int a = tid; a < L.n_in; a += nt) {
            double kc = B.kct[B.kidx[a]];
            if (thermo == THERMO_NH && a < L.n_ion)
              for (int q = 0; q < 3; q++) B.v[3 * a + q] = (B.v[3 * a + q] + B.f[3 * a + q] * kc) * fac;
            else
              for (int q = 0; q < 3; q++) B.v[3 * a + q] += B.f[3 * a + q] * kc;
          }
        }
        acc_steps += L.n_in;
        if (eflag) {
          for (int k = 0; k < U_N; k++) {
            double sv = wave_sum(e[k]);
            if ((tid & 63) == 0) B.red[(6 + k) * 16 + (tid >> 6)] = sv;
          }
          __syncthreads();
          if (tid < U_N) {
            double sv = 0;
            for (int w = 0; w < nt / 64; w++) sv += B.red[(6 + tid) * 16 + w];
            D->U[tid] = sv;
          }
        }
        __syncthreads();
      }
      if (!dead) {
        batch_store(R, L, B);
        if (tid < 3) { R.mol_x[tid] = B.sdbl[tid]; R.mol_v[tid] = B.sdbl[3 + tid]; }
        if (tid == 0) { D->loopPair = lp; D->zeta_ion = zeta; D->step = step_base + nsteps; }
      }
      if (tid == 0) {
        D->atom_steps += acc_steps;
        atomicMax(&D->shake_iters_max, B.sint[32]);
        if (B.sint[33]) atomicAdd(&D->shake_fail, (unsigned long long)B.sint[33]);
      }
    }
  report:
    // the counts the host plans the next launch with (thread 0 wrote them)
    if (tid == 0) cnt[order[rr]] = make_int4(D->n_in_gas, D->n_in_vap, D->pending, 0);
    __syncthreads();
  }
}

// ======================= host side =======================================
#define HIPCHK(x)                                                                              \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess) return set_error(MDN_E_HIP, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

struct HostReplica {
  SysDef def;
  std::vector<double> mol_x, mol_v, atom_x, atom_v, atom_f, atom_w;
  std::vector<uint8_t> region;
  std::vector<int> gi, gv, gg;  // reference format (global ids)
  int loopPair = 0, loop_update = 0;
  long long rebuild_count = 0;
  double U[U_N] = {0};
  uint64_t seed = 1;
  uint32_t replica_id = 0;
  int cap_gi = 0, cap_gv = 0, cap_gg = 0;
  std::vector<int> long_pairs;
  DevReplica dev;
  size_t off_dyn = 0;
  hipGraphExec_t graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [precision][eflag]
  int thermo = THERMO_NVE;  // ion thermostat (host copy of Dyn::thermo)
  double Tnh = 0, zeta_ion = 0;
  // diagnostics carried over a re-upload (mdn_set_state of any replica)
  long long atom_steps = 0, step = 0;
  int status = 0, shake_iters_max = 0;
  unsigned long long shake_fail = 0;
  // a step left unfinished by the batched kernel (Dyn::pending)
  int pending = 0;
};

}  // namespace mdn

struct mdn_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<mdn::HostReplica> reps;
  char *dbuf = nullptr;
  size_t dbytes = 0;
  mdn::DevReplica *d_reps = nullptr;
  bool committed = false;
  int batch_cap = 0, batch_threads = 256, batch_grid = 0;
  mdn::Dyn *d_dyn = nullptr;
  int last_cap = 0, last_lds = 0, batch_chunk = 1000, batch_split = 1, n_cu = 256;
  // split-launch tuning: workgroup sizes of the own-CU and shared-CU
  // classes, and the cost (a multiple of the median) above which a replica
  // may get a CU of its own
  int own_threads = 512, shared_threads = 128;
  double heavy_factor = 2.0;
  // direction and result buffers of mdn_force_jvp ([2][jvp_cap] doubles)
  double *d_jvp = nullptr;
  size_t jvp_cap = 0;
  hipStream_t stream2 = nullptr, stream3 = nullptr;
  hipEvent_t evj = nullptr, evk = nullptr, evl = nullptr;
  int *d_order = nullptr;
  double last_ms = 0;
  int last_heavy = 0;
  // launch plan of the last chunk of mdn_run_batched: replicas in the
  // own-CU, three-per-CU and four-per-CU classes of a split launch, and in a
  // uniform launch; the replica ids in launch order
  int last_classes[4] = {0, 0, 0, 0};
  std::vector<int> last_order;
  int any_shake = 0;
  int log_cap = 0;          // vapor events per replica between drains (0: off)
  double *d_obs = nullptr;  // [8 * n_rep] observables
  // calls that may change the replicas' device state count up `mut`; the
  // pinned counts h_dyn are those of the device while dyn_mut == mut (set
  // when mdn_run_batched leaves them current), so the next batched call
  // needs no read-back
  unsigned long long mut = 0, dyn_mut = ~0ull;
  int lds_attr = 0;  // k_batched variants whose LDS limit is raised
  bool plan_on_device = false;  // d_order holds plan.order
  // the split launch's classes, reused by the next chunk or call while they
  // are younger than a chunk, nothing but mdn_run_batched has touched the
  // replicas, the launch settings are the same and every replica's ion +
  // vapor and in-domain atom counts stay within kPlanSlack of those the
  // classes were made for (the capacities, sized per launch, keep their
  // headroom)
  struct SplitPlan {
    bool valid = false;
    long long age = 0;
    unsigned long long mut = 0;
    size_t rl = 0;
    int own_threads = 0, shared_threads = 0;
    double heavy_factor = 0;
    std::vector<int> heavy, clsB, clsA, order, need, need_h;
  } plan;
  // pinned host copies for the batched chunk loop: the replicas' counts
  // (read back behind each chunk's launches) and the launch order
  mdn::Dyn *h_dyn = nullptr;
  int *h_order = nullptr;
  // per replica {in-domain gas, in-domain vapor, pending}: written by the
  // batched kernel into pinned host memory as each replica's launch ends
  int4 *h_cnt = nullptr, *d_cnt = nullptr;
  int h_cap = 0;
};

namespace mdn {

// AMBER exclusion list, initialAMBER (compute_AMBER.cpp:256-298): ion pairs
// i<j that are not bonded, rigid, angle ends or dihedral ends
static std::vector<int> amber_long_pairs(const Species &I) {
  int n = I.natoms;
  std::vector<uint8_t> ex((size_t)n * n, 0);
  auto mark = [&](int a, int b) { ex[(size_t)a * n + b] = ex[(size_t)b * n + a] = 1; };
  for (size_t k = 0; 3 * k < I.bond.size(); k++) mark(I.bond[3 * k], I.bond[3 * k + 1]);
  for (size_t k = 0; 3 * k < I.rigid.size(); k++) mark(I.rigid[3 * k], I.rigid[3 * k + 1]);
  for (size_t k = 0; 4 * k < I.angle.size(); k++) mark(I.angle[4 * k], I.angle[4 * k + 2]);
  for (size_t k = 0; 5 * k < I.dih.size(); k++) mark(I.dih[5 * k], I.dih[5 * k + 3]);
  std::vector<int> out;
  for (int i = 0; i < n - 1; i++)
    for (int j = i + 1; j < n; j++)
      if (!ex[(size_t)i * n + j]) { out.push_back(i); out.push_back(j); }
  return out;
}

// bump layout over one device allocation; identical read-only tables are
// stored once (replicas of one batch usually share species)
struct Layout {
  std::vector<char> host;
  std::map<std::string, size_t> dedup;
  size_t put(const void *p, size_t n, bool share) {
    size_t off = (host.size() + 255) & ~(size_t)255;
    if (share && n) {
      std::string key(reinterpret_cast<const char *>(p), n);
      auto it = dedup.find(key);
      if (it != dedup.end()) return it->second;
      host.resize(off + n);
      memcpy(host.data() + off, p, n);
      dedup.emplace(std::move(key), off);
      return off;
    }
    host.resize(off + (n ? n : 8), 0);
    if (n && p) memcpy(host.data() + off, p, n);
    return off;
  }
  template <class T> size_t vec(const std::vector<T> &v, bool share) { return put(v.data(), v.size() * sizeof(T), share); }
  size_t zeros(size_t n) { return put(nullptr, n, false); }
};

template <class T> static T *at(char *base, size_t off) { return reinterpret_cast<T *>(base + off); }

static void in_lists(const HostReplica &h, std::vector<int> &ig, std::vector<int> &iv) {
  ig.clear(); iv.clear();
  for (int i = 1; i <= h.def.n_gas; i++) if (h.region[i]) ig.push_back(i);
  for (int i = 1 + h.def.n_gas; i < h.def.n_mol; i++) if (h.region[i]) iv.push_back(i);
}

// reference-format lists (global ids) -> in-list positions
static int lists_to_local(const HostReplica &h, std::vector<int> &gi, std::vector<int> &gv, std::vector<int> &gg) {
  std::vector<int> ig, iv;
  in_lists(h, ig, iv);
  std::vector<int> pos(h.def.n_mol, -1);
  for (size_t k = 0; k < ig.size(); k++) pos[ig[k]] = (int)k;
  for (size_t k = 0; k < iv.size(); k++) pos[iv[k]] = (int)k;
  auto conv = [&](const std::vector<int> &src, std::vector<int> &dst, bool second_is_mol) -> int {
    dst.resize(src.size());
    for (size_t k = 0; k < src.size(); k += 2) {
      int a = src[k], b = src[k + 1];
      if (a < 0 || a >= h.def.n_mol || pos[a] < 0) return -1;
      dst[k] = pos[a];
      if (second_is_mol) {
        if (b < 0 || b >= h.def.n_mol || pos[b] < 0) return -1;
        dst[k + 1] = pos[b];
      } else {
        dst[k + 1] = b;
      }
    }
    return 0;
  };
  if (conv(h.gi, gi, false) || conv(h.gv, gv, true) || conv(h.gg, gg, true)) return -1;
  return 0;
}

static void drop_graphs(mdn_ctx *c) {
  for (auto &h : c->reps)
    for (int p = 0; p < 2; p++)
      for (int e = 0; e < 2; e++)
        if (h.graph[p][e]) { (void)hipGraphExecDestroy(h.graph[p][e]); h.graph[p][e] = nullptr; }
}

// The rigid triples of a species ordered by constraint group (connected
// components of the constraint graph, in order of first appearance, merged
// into at most 4), each group keeping the file order; with D, also the
// group bounds and atom masks.  Constraints of different groups move
// different atoms, so a sweep in this order gives the same positions, to
// the bit, as one in the file order (shake.cpp:51-90).
static std::vector<int> shake_groups(const Species &S, DevSpecies *D) {
  const int nr = (int)S.rigid.size() / 3;
  std::vector<int> comp(S.natoms);
  for (int a = 0; a < S.natoms; a++) comp[a] = a;
  std::function<int(int)> find = [&](int a) { return comp[a] == a ? a : comp[a] = find(comp[a]); };
  for (int k = 0; k < nr; k++) {
    const int a = find(S.rigid[3 * k]), b = find(S.rigid[3 * k + 1]);
    if (a != b) comp[std::max(a, b)] = std::min(a, b);
  }
  std::vector<int> gid(nr), roots;
  for (int k = 0; k < nr; k++) {
    const int root = find(S.rigid[3 * k]);
    int g = (int)(std::find(roots.begin(), roots.end(), root) - roots.begin());
    if (g == (int)roots.size()) roots.push_back(root);
    gid[k] = S.natoms <= 64 ? std::min(g, 3) : 0;
  }
  const int ng = nr == 0 ? 1 : *std::max_element(gid.begin(), gid.end()) + 1;
  std::vector<int> out;
  int gs[5] = {0, 0, 0, 0, 0};
  unsigned long long gm[4] = {0, 0, 0, 0};
  for (int g = 0; g < ng; g++) {
    gs[g] = (int)out.size() / 3;
    for (int k = 0; k < nr; k++)
      if (gid[k] == g) {
        out.insert(out.end(), S.rigid.begin() + 3 * k, S.rigid.begin() + 3 * k + 3);
        for (int e = 0; e < 2; e++)
          if (S.rigid[3 * k + e] < 64) gm[g] |= 1ull << S.rigid[3 * k + e];
      }
  }
  gs[ng] = nr;
  // star groups: one centre atom joined to distinct satellites, at most
  // kStarMax constraints (in the group order of `out`)
  bool star = nr > 0;
  for (int g = 0; g < ng && star; g++) {
    const int k0 = gs[g], k1 = gs[g + 1], n = k1 - k0;
    if (n < 1 || n > kStarMax) { star = false; break; }
    const int i0 = out[3 * k0], j0 = out[3 * k0 + 1];
    const int c = n >= 2 && !(i0 == out[3 * k0 + 3] || i0 == out[3 * k0 + 4]) ? j0 : i0;
    std::vector<int> sats;
    for (int k = k0; k < k1; k++) {
      const int i = out[3 * k], j = out[3 * k + 1];
      if (i == j || (i != c && j != c)) { star = false; break; }
      const int sk = i == c ? j : i;
      if (std::find(sats.begin(), sats.end(), sk) != sats.end()) { star = false; break; }
      sats.push_back(sk);
    }
  }
  if (D) {
    D->shake_star = star ? 1 : 0;
    D->shake_ng = ng;
    for (int g = 0; g < 5; g++) D->shake_gs[g] = g <= ng ? gs[g] : nr;
    for (int g = 0; g < 4; g++) D->shake_gm[g] = gm[g];
  }
  return out;
}

static int upload_all(mdn_ctx *c) {
  drop_graphs(c);
  Layout lay;
  struct Offs {
    size_t sp[3][12];
    size_t pc, bt, ct, dc, dm, am;
    size_t mol_x, mol_v, region, ax, av, af, aw, axo, ing, inv, gi, gv, gg, dyn, vlog;
  };
  std::vector<Offs> offs(c->reps.size());
  c->any_shake = 0;
  // every replica's Dyn in one block: the host reads all of them with one copy
  size_t dyn_block = lay.zeros(sizeof(Dyn) * c->reps.size());
  for (size_t r = 0; r < c->reps.size(); r++) {
    HostReplica &h = c->reps[r];
    const SysDef &d = h.def;
    Offs &o = offs[r];
    for (int s = 0; s < 3; s++) {
      const Species &S = d.sp[s];
      std::vector<double> inv(S.mass.size());
      for (size_t k = 0; k < S.mass.size(); k++) inv[k] = 1.0 / S.mass[k];
      o.sp[s][0] = lay.vec(S.mass, true);
      o.sp[s][1] = lay.vec(inv, true);
      o.sp[s][2] = lay.vec(S.charge, true);
      o.sp[s][3] = lay.vec(S.tmpl_x, true);
      o.sp[s][4] = lay.vec(S.type, true);
      o.sp[s][5] = lay.vec(S.bond, true);
      o.sp[s][6] = lay.vec(S.angle, true);
      o.sp[s][7] = lay.vec(S.dih, true);
      o.sp[s][8] = lay.vec(shake_groups(S, nullptr), true);
      std::vector<double> se(S.type.size()), sg(S.type.size());
      for (size_t k = 0; k < S.type.size(); k++) {
        se[k] = sqrt(d.atypes[S.type[k]].eps);
        sg[k] = d.atypes[S.type[k]].sigma;
      }
      o.sp[s][9] = lay.vec(se, true);
      o.sp[s][10] = lay.vec(sg, true);
      if (d.shake && S.rigid.size()) c->any_shake = 1;
    }
    o.pc = lay.vec(d.pair_coeff, true);
    o.bt = lay.vec(d.btype, true);
    o.ct = lay.vec(d.ctype, true);
    o.dc = lay.vec(d.dcoeff, true);
    o.dm = lay.vec(d.dmulti, true);
    h.long_pairs = (d.flags & MDN_ION_AMBER) ? amber_long_pairs(d.sp[0]) : std::vector<int>();
    {
      int n = d.sp[0].natoms, W = (n + 31) / 32;
      std::vector<uint32_t> allow((size_t)std::max(1, n * W), 0u);
      for (size_t k = 0; k + 1 < h.long_pairs.size(); k += 2) {
        int i = h.long_pairs[k], j = h.long_pairs[k + 1];
        allow[(size_t)i * W + (j >> 5)] |= 1u << (j & 31);
        allow[(size_t)j * W + (i >> 5)] |= 1u << (i & 31);
      }
      o.am = lay.vec(allow, true);
    }
    o.mol_x = lay.vec(h.mol_x, false);
    o.mol_v = lay.vec(h.mol_v, false);
    o.region = lay.vec(h.region, false);
    o.ax = lay.vec(h.atom_x, false);
    o.av = lay.vec(h.atom_v, false);
    o.af = lay.vec(h.atom_f, false);
    o.aw = lay.vec(h.atom_w, false);
    o.axo = lay.zeros(d.shake ? sizeof(double) * 3 * (size_t)d.n_atoms : 8);
    std::vector<int> ig, iv, lgi, lgv, lgg;
    in_lists(h, ig, iv);
    if (lists_to_local(h, lgi, lgv, lgg)) return set_error(MDN_E_ARG, "state pair list names a molecule that is not in the domain");
    ig.resize(std::max(d.n_gas, 1), 0);
    iv.resize(std::max(d.n_vap, 1), 0);
    o.ing = lay.vec(ig, false);
    o.inv = lay.vec(iv, false);
    lgi.resize(2 * (size_t)std::max(h.cap_gi, 1), 0);
    lgv.resize(2 * (size_t)std::max(h.cap_gv, 1), 0);
    lgg.resize(2 * (size_t)std::max(h.cap_gg, 1), 0);
    o.gi = lay.vec(lgi, false);
    o.gv = lay.vec(lgv, false);
    o.gg = lay.vec(lgg, false);
    o.vlog = c->log_cap > 0 ? lay.zeros(sizeof(VapEvent) * (size_t)c->log_cap) : 0;
    Dyn dyn;
    memset(&dyn, 0, sizeof(dyn));
    dyn.atom_steps = h.atom_steps;
    dyn.step = h.step;
    dyn.status = h.status;
    dyn.shake_iters_max = h.shake_iters_max;
    dyn.shake_fail = h.shake_fail;
    dyn.pending = h.pending;
    for (int k = 0; k < U_N; k++) dyn.U[k] = h.U[k];
    dyn.rebuild_count = h.rebuild_count;
    dyn.loopPair = h.loopPair;
    dyn.loop_update = h.loop_update;
    dyn.n_gi = (int)h.gi.size() / 2;
    dyn.n_gv = (int)h.gv.size() / 2;
    dyn.n_gg = (int)h.gg.size() / 2;
    std::vector<int> ig2, iv2;
    in_lists(h, ig2, iv2);
    dyn.n_in_gas = (int)ig2.size();
    dyn.n_in_vap = (int)iv2.size();
    dyn.thermo = h.thermo;
    dyn.zeta_ion = h.zeta_ion;
    dyn.Tnh_ion = h.Tnh;
    o.dyn = dyn_block + sizeof(Dyn) * r;
    memcpy(lay.host.data() + o.dyn, &dyn, sizeof(dyn));
    h.off_dyn = o.dyn;
  }
  size_t reps_off = lay.zeros(sizeof(DevReplica) * c->reps.size());
  size_t order_off = lay.zeros(sizeof(int) * c->reps.size());
  size_t obs_off = lay.zeros(sizeof(double) * 8 * c->reps.size());
  if (c->dbuf) { hipFree(c->dbuf); c->dbuf = nullptr; }
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMalloc(&c->dbuf, lay.host.size()));
  c->dbytes = lay.host.size();
  char *B = c->dbuf;
  std::vector<DevReplica> drs(c->reps.size());
  for (size_t r = 0; r < c->reps.size(); r++) {
    HostReplica &h = c->reps[r];
    const SysDef &d = h.def;
    Offs &o = offs[r];
    DevReplica &R = drs[r];
    memset(&R, 0, sizeof(R));
    for (int s = 0; s < 3; s++) {
      const Species &S = d.sp[s];
      DevSpecies &D = R.sp[s];
      D.natoms = S.natoms;
      D.nbond = (int)S.bond.size() / 3; D.nangle = (int)S.angle.size() / 4;
      D.ndih = (int)S.dih.size() / 5; D.nrigid = d.shake ? (int)S.rigid.size() / 3 : 0;
      D.molmass = S.molmass;
      D.mass = at<double>(B, o.sp[s][0]); D.invmass = at<double>(B, o.sp[s][1]);
      D.charge = at<double>(B, o.sp[s][2]); D.tmpl = at<double>(B, o.sp[s][3]);
      D.type = at<int>(B, o.sp[s][4]); D.bond = at<int>(B, o.sp[s][5]); D.angle = at<int>(B, o.sp[s][6]);
      D.dih = at<int>(B, o.sp[s][7]); D.rigid = at<int>(B, o.sp[s][8]);
      shake_groups(S, &D);
      D.se = at<double>(B, o.sp[s][9]); D.sg = at<double>(B, o.sp[s][10]);
    }
    R.n_gas = d.n_gas; R.n_vap = d.n_vap; R.n_mol = d.n_mol; R.n_atoms = d.n_atoms;
    R.n_types = (int)d.atypes.size();
    R.off_gas = d.sp[0].natoms;
    R.off_vap = d.sp[0].natoms + d.n_gas * d.sp[1].natoms;
    R.flags = d.flags;
    R.shake = d.shake; R.shake_max_iter = d.shake_max_iter; R.gas_spin = d.gas_spin;
    R.ion_type_base = d.ion_type_base;
    R.dt = d.dt; R.d_size = d.d_size; R.margin = d.margin; R.ML2 = d.ML2;
    R.qqrd2e = 1.602e-19 * 1.602e-19 / 4.0 / M_PI / 8.85e-12 * 6.02e23 * 1e10 / 4184.0;  // constants.hpp:50
    R.sig_gas = sqrt(1.38e-23 * d.T / d.m_gas_kg);
    R.sig_vap = d.m_vap_kg > 0 ? sqrt(1.38e-23 * d.T / d.m_vap_kg) : 0.0;
    for (int k = 0; k < 3; k++) R.efield[k] = d.efield[k];
    R.key0 = (uint32_t)h.seed;
    R.key1 = (uint32_t)(h.seed >> 32) ^ h.replica_id;
    R.pair_coeff = at<double>(B, o.pc); R.btype = at<double>(B, o.bt); R.ctype = at<double>(B, o.ct);
    R.dcoeff = at<double>(B, o.dc); R.dmulti = at<int>(B, o.dm);
    R.amber_words = (d.sp[0].natoms + 31) / 32; R.amber_allow = at<uint32_t>(B, o.am);
    R.mol_x = at<double>(B, o.mol_x); R.mol_v = at<double>(B, o.mol_v); R.region = at<uint8_t>(B, o.region);
    R.atom_x = at<double>(B, o.ax); R.atom_v = at<double>(B, o.av); R.atom_f = at<double>(B, o.af);
    R.atom_w = at<double>(B, o.aw); R.atom_xo = at<double>(B, o.axo);
    R.in_gas = at<int>(B, o.ing); R.in_vap = at<int>(B, o.inv);
    R.gi = at<int>(B, o.gi); R.gv = at<int>(B, o.gv); R.gg = at<int>(B, o.gg);
    R.cap_gi = h.cap_gi; R.cap_gv = h.cap_gv; R.cap_gg = h.cap_gg;
    R.dyn = at<Dyn>(B, o.dyn);
    R.vlog = c->log_cap > 0 ? at<VapEvent>(B, o.vlog) : nullptr;
    R.log_cap = c->log_cap;
    h.dev = R;
  }
  memcpy(lay.host.data() + reps_off, drs.data(), sizeof(DevReplica) * drs.size());
  c->d_reps = at<DevReplica>(B, reps_off);
  c->d_dyn = at<Dyn>(B, dyn_block);
  c->d_order = at<int>(B, order_off);
  c->plan_on_device = false;
  c->d_obs = at<double>(B, obs_off);
  HIPCHK(hipMemcpy(c->dbuf, lay.host.data(), lay.host.size(), hipMemcpyHostToDevice));
  return MDN_OK;
}

static int status_to_code(int st) {
  if (st & (ST_CAP_GI | ST_CAP_GV | ST_CAP_GG | ST_LDS_CAP)) return MDN_E_CAPACITY;
  return MDN_OK;
}

static int check_replica(mdn_ctx *c, int32_t r) {
  if (!c) return set_error(MDN_E_ARG, "null context");
  if (!c->committed) return set_error(MDN_E_STATE, "call mdn_commit first");
  if (r < 0 || r >= (int)c->reps.size()) return set_error(MDN_E_ARG, "replica index out of range");
  return MDN_OK;
}

static int grid_for(int n, int blk = kBlk, int cap = 2048) {
  int g = (n + blk - 1) / blk;
  return std::max(1, std::min(g, cap));
}
// k_com_weights: a 1024-thread workgroup per 1024 atoms of the replica (an
// upper bound of the in-domain count), at most 32
static int com_blocks(int natoms) { return std::max(1, std::min(32, (natoms + 1023) / 1024)); }

// every selected potential of one replica through the per-call kernels; vv
// goes to an all-pairs kernel when the vapor's types fit its table: the
// Newton-3 one when the vapor (in-domain or not) holds at least
// kVVNewtonMin atoms, the two-sided tiled one below that
constexpr int kVVNewtonMin = 4096;
template <typename Rl>
static int launch_terms(mdn_ctx *c, const HostReplica &h, const DevReplica *R, uint32_t bits, bool eflag) {
  const SysDef &d = h.def;
  int tv0 = d.sp[2].natoms ? *std::min_element(d.sp[2].type.begin(), d.sp[2].type.end()) : 0;
  int ntv = d.sp[2].natoms ? *std::max_element(d.sp[2].type.begin(), d.sp[2].type.end()) - tv0 + 1 : 1;
  bool tiled = ntv <= 8;
  uint32_t rest = tiled ? bits & ~(uint32_t)MDN_VV : bits;
  const int nva_b = d.n_vap * d.sp[2].natoms;  // bound on the in-domain vapor atoms
  if (rest && (bits & MDN_VV) && tiled && nva_b < kVVNewtonMin) {
    const int gx = std::max(1, (nva_b + kVVTile - 1) / kVVTile);
    const int gy = std::max(1, std::min(32, (4 * c->n_cu + gx - 1) / gx));
    if (eflag)
      hipLaunchKernelGGL((k_terms_vv<Rl, true>), dim3(kTermsGrid + gx * gy), dim3(kBlk), 0, c->stream, R, rest, tv0, ntv,
                         gx, gy);
    else
      hipLaunchKernelGGL((k_terms_vv<Rl, false>), dim3(kTermsGrid + gx * gy), dim3(kBlk), 0, c->stream, R, rest, tv0,
                         ntv, gx, gy);
    HIPCHK(hipGetLastError());
    return MDN_OK;
  }
  if (rest) hipLaunchKernelGGL(k_terms<Rl>, dim3(kTermsGrid), dim3(kBlk), 0, c->stream, R, rest, eflag ? 1 : 0);
  if ((bits & MDN_VV) && tiled) {
    const int nva = d.n_vap * d.sp[2].natoms;  // bound on the in-domain vapor atoms
    if (nva >= kVVNewtonMin) {
      // one workgroup per super-block pair, for the largest possible
      // in-domain vapor set
      const int nsb = std::max(1, (nva + kVVSuper - 1) / kVVSuper);
      const int gx = nsb * (nsb + 1) / 2;
      if (eflag) hipLaunchKernelGGL((k_vv_n3<Rl, true>), dim3(gx), dim3(kVVSuper), 0, c->stream, R, tv0, ntv);
      else hipLaunchKernelGGL((k_vv_n3<Rl, false>), dim3(gx), dim3(kVVSuper), 0, c->stream, R, tv0, ntv);
    } else {
      // the partner range split gy ways for about four workgroups per CU,
      // at most 32: each split adds one FP64 atomic per atom and component
      // onto the same force slots (for a 600-atom vapor set, 32 splits
      // instead of 64 shortened a 100-step segment gradient by 3 %)
      const int gx = std::max(1, (nva + kVVTile - 1) / kVVTile);
      const int gy = std::max(1, std::min(32, (4 * c->n_cu + gx - 1) / gx));
      if (eflag) hipLaunchKernelGGL((k_vv_tiled<Rl, true>), dim3(gx, gy), dim3(kVVTile), 0, c->stream, R, tv0, ntv);
      else hipLaunchKernelGGL((k_vv_tiled<Rl, false>), dim3(gx, gy), dim3(kVVTile), 0, c->stream, R, tv0, ntv);
    }
  }
  HIPCHK(hipGetLastError());
  return MDN_OK;
}

// the derivative along u of the forces of the selected pair and bonded terms
// (gi, vi, gg, gv, vv, AMBER, vapor and gas intra) at the replica's
// positions, in FP64, added to hu; the Efield force does not depend on the
// positions
__global__ void __launch_bounds__(kBlk) k_force_jvp(const DevReplica *__restrict__ Rp, uint32_t bits,
                                                    const double *__restrict__ u, double *__restrict__ hu) {
  const DevReplica &R = *Rp;
  JStore s;
  s.R = Rp;
  s.L = view_from_dyn(R);
  s.u = u;
  s.hu = hu;
  const int t = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
  double e[U_N];
  // each term's items start where the previous term's ended (modulo the
  // grid, in whole waves), so that the terms run side by side rather than
  // one after another on the first waves
  const LView &L = s.L;
  int off = 0;
  auto at = [&](int items) {
    const int tt = t - off;
    off = (int)((off + ((long long)items + 63) / 64 * 64) % nt);
    return tt < 0 ? tt + nt : tt;
  };
  const DevSpecies &I = R.sp[0], &G = R.sp[1], &V = R.sp[2];
  const int nj = L.na_g < L.na_v ? L.na_g : L.na_v;
  if (bits & B_GI) term_gi<Dual>(R, s, at(L.n_gi * L.na_g), nt, false, e);
  if (bits & B_VI) term_vi<Dual>(R, s, at(L.nv * L.na_v * L.n_ion), nt, false, e);
  if (bits & B_GG) term_gg<Dual>(R, s, at(L.n_gg * L.na_g * L.na_g), nt, false, e);
  if (bits & B_GV) term_gv<Dual>(R, s, at(L.n_gv * L.na_g * nj), nt, false, e);
  if (bits & B_VV) {
    const int tv = at(64 * L.nv * L.na_v);
    term_vv_jvp_rows(R, s, tv >> 6, nt >> 6, tv & 63);
  }
  if (bits & B_AMBER) {
    const int ta = at(64 * I.natoms);
    term_amber_long_rows<Dual>(R, s, ta >> 6, nt >> 6, ta & 63, false, e);
    term_amber_bonded(R, s, at(I.nbond + I.nangle + I.ndih), nt, false, e);
  }
  if (bits & B_VINTRA) term_vintra(R, s, at(L.nv * (V.nbond + V.nangle + V.ndih)), nt, false, e);
  if (bits & B_GINTRA) term_gintra(R, s, at(L.ng * G.nbond), nt, false, e);
}

// start of the rest of a step the batched kernel left unfinished: clears the
// flag and, on an energy step, the energy buckets
__global__ void k_resume(const DevReplica *__restrict__ Rp, int zeroU) {
  const DevReplica &R = *Rp;
  if (threadIdx.x == 0) {
    if (zeroU)
      for (int k = 0; k < U_N; k++) R.dyn->U[k] = 0;
    R.dyn->pending = 0;
  }
}

// the rest of that step: the forces (verlet.cpp:31-32) and the second
// half-kick (verlet.cpp:34, 36)
template <typename Rl>
static int launch_resume(mdn_ctx *c, HostReplica &h, const DevReplica *R, bool eflag) {
  const SysDef &d = h.def;
  hipLaunchKernelGGL(k_resume, dim3(1), dim3(64), 0, c->stream, R, eflag ? 1 : 0);
  int rc = launch_terms<Rl>(c, h, R, d.flags, eflag);
  if (rc) return rc;
  hipLaunchKernelGGL(k_kick, dim3(grid_for(d.n_atoms)), dim3(kBlk), 0, c->stream, R, 0, 1,
                     h.thermo == THERMO_NH ? 1 : 0);
  HIPCHK(hipGetLastError());
  h.pending = 0;
  return MDN_OK;
}

// MD::shake (shake.cpp:44-92) of every constrained molecule: the vapor a
// lane per constraint group when its groups are stars, the rest (and any
// other vapor) one thread per molecule
static void launch_shake(mdn_ctx *c, const HostReplica &h, const DevReplica *R) {
  const SysDef &d = h.def;
  const DevSpecies &V = h.dev.sp[2];
  const bool grp = V.nrigid > 0 && V.nrigid <= kShakeGrpCap && V.shake_star;
  if (grp) {
    const int W = V.shake_ng >= 4 ? 4 : V.shake_ng >= 2 ? 2 : 1;
    hipLaunchKernelGGL(k_shake_groups, dim3(grid_for(d.n_vap * W, kShakeGrpBlk)), dim3(kShakeGrpBlk), 0, c->stream, R);
  }
  if (!grp || h.dev.sp[0].nrigid || h.dev.sp[1].nrigid)
    hipLaunchKernelGGL(k_shake, dim3(grid_for(d.n_mol, kShakeBlk)), dim3(kShakeBlk), 0, c->stream, R, grp ? 1 : 0);
}

// one MD::verlet step of one replica through the per-call kernels
template <typename Rl>
static int launch_step(mdn_ctx *c, const HostReplica &h, const DevReplica *R, bool eflag) {
  const SysDef &d = h.def;
  int g = grid_for(d.n_atoms);
  const int nh = h.thermo == THERMO_NH ? 1 : 0;
  hipLaunchKernelGGL(k_kick, dim3(g), dim3(kBlk), 0, c->stream, R, eflag ? 1 : 0, 0, 0);
  // thermostat hooks where MD::verlet calls them (verlet.cpp:22, 27, 36)
  if (h.thermo == THERMO_SCALE) hipLaunchKernelGGL(k_ion_scale, dim3(1), dim3(1024), 0, c->stream, R);
  bool sh = d.shake && (d.sp[0].rigid.size() || d.sp[1].rigid.size() || d.sp[2].rigid.size());
  hipLaunchKernelGGL(k_drift, dim3(g), dim3(kBlk), 0, c->stream, R, sh ? 1 : 0);
  if (sh) launch_shake(c, h, R);
  hipLaunchKernelGGL(k_com_weights, dim3(com_blocks(d.n_atoms)), dim3(1024), 0, c->stream, R, nh);
  hipLaunchKernelGGL(k_rebuild, dim3(1), dim3(1024), 0, c->stream, R, 0);
  int rc = launch_terms<Rl>(c, h, R, d.flags, eflag);
  if (rc) return rc;
  hipLaunchKernelGGL(k_kick, dim3(g), dim3(kBlk), 0, c->stream, R, 0, 1, nh);
  HIPCHK(hipGetLastError());
  return MDN_OK;
}

}  // namespace mdn

using namespace mdn;

extern "C" int mdn_get_state(mdn_ctx *c, int32_t r, mdn_state *st);

namespace mdn {
// the device's state and counters of replica r into its host copy (the input
// of the next upload); with state == false only the counters
static int pull_replica(mdn_ctx *c, int32_t r, bool state) {
  HostReplica &h = c->reps[r];
  const DevReplica &R = h.dev;
  Dyn dyn;
  HIPCHK(hipMemcpy(&dyn, R.dyn, sizeof(Dyn), hipMemcpyDeviceToHost));
  h.atom_steps = dyn.atom_steps;
  h.step = dyn.step;
  h.status = dyn.status;
  h.shake_iters_max = dyn.shake_iters_max;
  h.shake_fail = dyn.shake_fail;
  h.pending = dyn.pending;
  if (!state) return MDN_OK;
  std::vector<int> gi(2 * (size_t)std::max(h.cap_gi, 1)), gv(2 * (size_t)std::max(h.cap_gv, 1)),
      gg(2 * (size_t)std::max(h.cap_gg, 1));
  mdn_state st;
  memset(&st, 0, sizeof(st));
  st.mol_x = h.mol_x.data(); st.mol_v = h.mol_v.data(); st.region = h.region.data();
  st.atom_x = h.atom_x.data(); st.atom_v = h.atom_v.data(); st.atom_f = h.atom_f.data();
  st.atom_w = h.atom_w.data();
  st.cap_gi = h.cap_gi; st.cap_gv = h.cap_gv; st.cap_gg = h.cap_gg;
  st.gi = gi.data(); st.gv = gv.data(); st.gg = gg.data();
  int rc = mdn_get_state(c, r, &st);
  if (rc && !(rc == MDN_E_CAPACITY && h.status)) return rc;  // a flagged replica keeps its flags
  h.loopPair = st.loopPair;
  h.loop_update = st.loop_update;
  h.rebuild_count = st.rebuild_count;
  h.zeta_ion = st.zeta_ion;
  for (int k = 0; k < U_N; k++) h.U[k] = st.U[k];
  h.gi.assign(gi.begin(), gi.begin() + 2 * (size_t)st.n_gi);
  h.gv.assign(gv.begin(), gv.begin() + 2 * (size_t)st.n_gv);
  h.gg.assign(gg.begin(), gg.begin() + 2 * (size_t)st.n_gg);
  return MDN_OK;
}
}  // namespace mdn

extern "C" {

const char *mdn_last_error(void) { return g_err.c_str(); }

int mdn_create(int device, mdn_ctx **out) {
  if (!out) return set_error(MDN_E_ARG, "mdn_create: null out");
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return set_error(MDN_E_ARG, "mdn_create: no such device");
  mdn_ctx *c = new mdn_ctx();
  c->device = device;
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&c->ev0));
  HIPCHK(hipEventCreate(&c->ev1));
  HIPCHK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&c->stream3, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&c->evj, hipEventDisableTiming));
  HIPCHK(hipEventCreate(&c->evk));
  HIPCHK(hipEventCreate(&c->evl));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  c->n_cu = prop.multiProcessorCount;
  *out = c;
  return MDN_OK;
}

void mdn_destroy(mdn_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  drop_graphs(c);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->dbuf) hipFree(c->dbuf);
  if (c->d_jvp) hipFree(c->d_jvp);
  if (c->h_dyn) hipHostFree(c->h_dyn);
  if (c->h_order) hipHostFree(c->h_order);
  if (c->h_cnt) hipHostFree(c->h_cnt);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->evj) hipEventDestroy(c->evj);
  if (c->evk) hipEventDestroy(c->evk);
  if (c->evl) hipEventDestroy(c->evl);
  if (c->stream2) { hipStreamSynchronize(c->stream2); hipStreamDestroy(c->stream2); }
  if (c->stream3) { hipStreamSynchronize(c->stream3); hipStreamDestroy(c->stream3); }
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

static int copy_state_in(HostReplica &h, const mdn_state *st) {
  const SysDef &d = h.def;
  int nm = d.n_mol, na = d.n_atoms;
  if (!st->mol_x || !st->mol_v || !st->region || !st->atom_x || !st->atom_v || !st->atom_f || !st->atom_w)
    return set_error(MDN_E_ARG, "state: null array");
  h.mol_x.assign(st->mol_x, st->mol_x + 3 * (size_t)nm);
  h.mol_v.assign(st->mol_v, st->mol_v + 3 * (size_t)nm);
  h.region.assign(st->region, st->region + nm);
  h.atom_x.assign(st->atom_x, st->atom_x + 3 * (size_t)na);
  h.atom_v.assign(st->atom_v, st->atom_v + 3 * (size_t)na);
  h.atom_f.assign(st->atom_f, st->atom_f + 3 * (size_t)na);
  h.atom_w.assign(st->atom_w, st->atom_w + na);
  h.region[0] = 1;
  h.loopPair = st->loopPair;
  h.loop_update = st->loop_update;
  h.rebuild_count = st->rebuild_count;
  h.zeta_ion = st->zeta_ion;
  for (int k = 0; k < U_N; k++) h.U[k] = st->U[k];
  h.gi.assign(st->gi ? st->gi : nullptr, st->gi ? st->gi + 2 * (size_t)st->n_gi : nullptr);
  h.gv.assign(st->gv ? st->gv : nullptr, st->gv ? st->gv + 2 * (size_t)st->n_gv : nullptr);
  h.gg.assign(st->gg ? st->gg : nullptr, st->gg ? st->gg + 2 * (size_t)st->n_gg : nullptr);
  // list capacities: the caller's, else an estimate from the gas density in
  // the r < 110 A sphere (x3 + 64 molecules of headroom)
  double vol_dom = 4.0 / 3.0 * M_PI * 110.0 * 110.0 * 110.0;
  double frac = d.d_size > 0 ? vol_dom / (d.d_size * d.d_size * d.d_size) : 1.0;
  int g_est = (int)std::min<double>(d.n_gas, 3.0 * d.n_gas * std::min(1.0, frac) + 64);
  int v_est = (int)std::min<double>(d.n_vap, 3.0 * d.n_vap * std::min(1.0, frac) + 64);
  h.cap_gi = st->cap_gi > 0 ? st->cap_gi : std::max(64, g_est * d.sp[0].natoms);
  h.cap_gv = st->cap_gv > 0 ? st->cap_gv : std::max(64, g_est * std::max(1, v_est));
  h.cap_gg = st->cap_gg > 0 ? st->cap_gg : std::max(64, g_est * 32);
  h.cap_gi = std::max<int>(h.cap_gi, (int)h.gi.size() / 2);
  h.cap_gv = std::max<int>(h.cap_gv, (int)h.gv.size() / 2);
  h.cap_gg = std::max<int>(h.cap_gg, (int)h.gg.size() / 2);
  return MDN_OK;
}

int mdn_add_replica(mdn_ctx *c, const mdn_sysdef *sd, const mdn_state *st, uint64_t seed, uint32_t replica_id,
                    int32_t *handle) {
  if (c) c->mut++;
  if (!c || !sd || !st) return set_error(MDN_E_ARG, "mdn_add_replica: null argument");
  c->reps.emplace_back();
  HostReplica &h = c->reps.back();
  h.def = *reinterpret_cast<const SysDef *>(sd);
  h.seed = seed;
  h.replica_id = replica_id;
  // the condition file's NVTion line (until mdn_set_thermostat changes it)
  h.thermo = h.def.nvt_ion_mode;
  h.Tnh = h.def.Tnh_ion;
  int rc = copy_state_in(h, st);
  if (rc) { c->reps.pop_back(); return rc; }
  c->committed = false;
  if (handle) *handle = (int32_t)c->reps.size() - 1;
  return MDN_OK;
}

int mdn_commit(mdn_ctx *c) {
  if (c) c->mut++;
  if (!c) return set_error(MDN_E_ARG, "mdn_commit: null context");
  if (c->reps.empty()) return set_error(MDN_E_STATE, "mdn_commit: no replicas");
  int rc = upload_all(c);
  if (rc) return rc;
  c->committed = true;
  return MDN_OK;
}

int mdn_set_state(mdn_ctx *c, int32_t r, const mdn_state *st) {
  if (c) c->mut++;
  int rc = check_replica(c, r);
  if (rc) return rc;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  // the device holds every replica's current state: bring the others' (and
  // every replica's counters) back to the host copies the re-upload reads
  if (c->committed)
    for (int32_t q = 0; q < (int32_t)c->reps.size(); q++)
      if ((rc = pull_replica(c, q, q != r))) return rc;
  rc = copy_state_in(c->reps[r], st);
  if (rc) return rc;
  // a new state starts at a step boundary with no capacity flags
  c->reps[r].pending = 0;
  c->reps[r].status = 0;
  return mdn_commit(c);
}

// MD::setNVE (nvt.cpp:38-43) and the NVTion modes (setCondition.cpp:158-172)
// for one replica (r >= 0) or all (r = -1); Nose-Hoover also resets zeta
int mdn_set_thermostat(mdn_ctx *c, int32_t r, int32_t mode, double Tnh) {
  if (c) c->mut++;
  if (!c) return set_error(MDN_E_ARG, "mdn_set_thermostat: null context");
  if (mode != MDN_NVE && mode != MDN_VSCALE && mode != MDN_NOSE_HOOVER)
    return set_error(MDN_E_ARG, "mdn_set_thermostat: mode must be MDN_NVE, MDN_VSCALE or MDN_NOSE_HOOVER");
  if (mode == MDN_NOSE_HOOVER && !(Tnh >= 0)) return set_error(MDN_E_ARG, "mdn_set_thermostat: Tnh must be >= 0");
  int r0 = r, r1 = r + 1;
  if (r == -1) { r0 = 0; r1 = (int)c->reps.size(); }
  else if (int rc = check_replica(c, r)) return rc;
  HIPCHK(hipSetDevice(c->device));
  drop_graphs(c);
  for (int q = r0; q < r1; q++) {
    HostReplica &h = c->reps[q];
    h.thermo = mode;
    if (mode == MDN_NOSE_HOOVER) { h.Tnh = Tnh; h.zeta_ion = 0; }
    if (c->committed) {
      // read-modify-write of the replica's Dyn (every stream has joined
      // c->stream, so no kernel holds it)
      Dyn dyn;
      HIPCHK(hipStreamSynchronize(c->stream));
      HIPCHK(hipMemcpy(&dyn, h.dev.dyn, sizeof(Dyn), hipMemcpyDeviceToHost));
      dyn.thermo = mode;
      if (mode == MDN_NOSE_HOOVER) { dyn.zeta_ion = 0; dyn.Tnh_ion = Tnh; }
      HIPCHK(hipMemcpy(h.dev.dyn, &dyn, sizeof(Dyn), hipMemcpyHostToDevice));
    }
  }
  return MDN_OK;
}

}  // extern "C"

namespace mdn {
int engine_replicas(const mdn_ctx *c) { return c && c->committed ? (int)c->reps.size() : -1; }
const SysDef &engine_sysdef(const mdn_ctx *c, int r) { return c->reps[r].def; }
int engine_log_cap(const mdn_ctx *c) { return c->log_cap; }
}  // namespace mdn

extern "C" {

static_assert(sizeof(VapEvent) == sizeof(mdn_vapor_event) && offsetof(VapEvent, dv) == offsetof(mdn_vapor_event, dv),
              "VapEvent mirrors mdn_vapor_event");

// the replica range of an entry point that takes one replica or all (-1)
static int replica_range(mdn_ctx *c, int32_t r, int *r0, int *r1) {
  if (!c) return set_error(MDN_E_ARG, "null context");
  if (!c->committed) return set_error(MDN_E_STATE, "call mdn_commit first");
  if (r == -1) { *r0 = 0; *r1 = (int)c->reps.size(); return MDN_OK; }
  if (int rc = check_replica(c, r)) return rc;
  *r0 = r; *r1 = r + 1;
  return MDN_OK;
}

int mdn_set_vapor_log(mdn_ctx *c, int32_t cap) {
  if (c) c->mut++;
  if (!c) return set_error(MDN_E_ARG, "mdn_set_vapor_log: null context");
  if (cap < 0) return set_error(MDN_E_ARG, "mdn_set_vapor_log: negative capacity");
  if (cap == c->log_cap) return MDN_OK;
  int rc;
  if (c->committed) {
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int32_t q = 0; q < (int32_t)c->reps.size(); q++)
      if ((rc = pull_replica(c, q, true))) return rc;
  }
  c->log_cap = cap;
  return c->committed ? mdn_commit(c) : MDN_OK;
}

int mdn_drain_vapor_log(mdn_ctx *c, int32_t r, mdn_vapor_event *out, int32_t cap, int32_t *n) {
  if (c) c->mut++;
  int rc = check_replica(c, r);
  if (rc) return rc;
  if (!n || (cap > 0 && !out)) return set_error(MDN_E_ARG, "mdn_drain_vapor_log: null argument");
  *n = 0;
  if (c->log_cap == 0) return MDN_OK;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  const DevReplica &R = c->reps[r].dev;
  char *nlog = reinterpret_cast<char *>(R.dyn) + offsetof(Dyn, n_log);
  int recorded = 0;
  HIPCHK(hipMemcpy(&recorded, nlog, sizeof(int), hipMemcpyDeviceToHost));
  const int kept = std::min(recorded, c->log_cap);
  if (kept > cap) return set_error(MDN_E_CAPACITY, "mdn_drain_vapor_log: output buffer holds fewer events than recorded");
  if (kept) HIPCHK(hipMemcpy(out, R.vlog, sizeof(VapEvent) * (size_t)kept, hipMemcpyDeviceToHost));
  const int zero = 0;
  HIPCHK(hipMemcpy(nlog, &zero, sizeof(int), hipMemcpyHostToDevice));
  *n = kept;
  if (recorded > kept)
    return set_error(MDN_E_CAPACITY, "mdn_drain_vapor_log: " + std::to_string(recorded - kept) +
                                         " events dropped (log capacity " + std::to_string(c->log_cap) + ")");
  return MDN_OK;
}

int mdn_observe(mdn_ctx *c, int32_t r, double *out) {
  int r0, r1;
  int rc = replica_range(c, r, &r0, &r1);
  if (rc) return rc;
  if (!out) return set_error(MDN_E_ARG, "mdn_observe: null output");
  HIPCHK(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_observe, dim3(r1 - r0), dim3(256), 0, c->stream, c->d_reps, r0, c->d_obs);
  HIPCHK(hipGetLastError());
  std::vector<double> o(8 * (size_t)(r1 - r0));
  HIPCHK(hipMemcpyAsync(o.data(), c->d_obs, sizeof(double) * o.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int q = 0; q < r1 - r0; q++)
    for (int k = 0; k < 6; k++) out[6 * (size_t)q + k] = o[8 * (size_t)q + k];
  return MDN_OK;
}

int mdn_set_step(mdn_ctx *c, int32_t r, int64_t step) {
  if (c) c->mut++;
  int r0, r1;
  int rc = replica_range(c, r, &r0, &r1);
  if (rc) return rc;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  const long long v = step;
  for (int q = r0; q < r1; q++)
    HIPCHK(hipMemcpy(reinterpret_cast<char *>(c->reps[q].dev.dyn) + offsetof(Dyn, step), &v, sizeof(v),
                     hipMemcpyHostToDevice));
  return MDN_OK;
}

// run.cpp:42-53 as written: the loop zeroes position[0] on its first pass and
// then subtracts zero from every other position
int mdn_recentre(mdn_ctx *c, int32_t r) {
  if (c) c->mut++;
  int r0, r1;
  int rc = replica_range(c, r, &r0, &r1);
  if (rc) return rc;
  HIPCHK(hipSetDevice(c->device));
  for (int q = r0; q < r1; q++) HIPCHK(hipMemsetAsync(c->reps[q].dev.mol_x, 0, 3 * sizeof(double), c->stream));
  return MDN_OK;
}

int mdn_get_state(mdn_ctx *c, int32_t r, mdn_state *st) {
  int rc = check_replica(c, r);
  if (rc) return rc;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  HostReplica &h = c->reps[r];
  const DevReplica &R = h.dev;
  const SysDef &d = h.def;
  int nm = d.n_mol, na = d.n_atoms;
  HIPCHK(hipMemcpy(st->mol_x, R.mol_x, sizeof(double) * 3 * nm, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(st->mol_v, R.mol_v, sizeof(double) * 3 * nm, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(st->region, R.region, nm, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(st->atom_x, R.atom_x, sizeof(double) * 3 * na, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(st->atom_v, R.atom_v, sizeof(double) * 3 * na, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(st->atom_f, R.atom_f, sizeof(double) * 3 * na, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(st->atom_w, R.atom_w, sizeof(double) * na, hipMemcpyDeviceToHost));
  Dyn dyn;
  HIPCHK(hipMemcpy(&dyn, R.dyn, sizeof(Dyn), hipMemcpyDeviceToHost));
  st->loopPair = dyn.loopPair;
  st->loop_update = dyn.loop_update;
  st->rebuild_count = dyn.rebuild_count;
  st->zeta_ion = dyn.zeta_ion;
  for (int k = 0; k < U_N; k++) st->U[k] = dyn.U[k];
  std::vector<int> ig(std::max(dyn.n_in_gas, 1)), iv(std::max(dyn.n_in_vap, 1));
  if (dyn.n_in_gas) HIPCHK(hipMemcpy(ig.data(), R.in_gas, sizeof(int) * dyn.n_in_gas, hipMemcpyDeviceToHost));
  if (dyn.n_in_vap) HIPCHK(hipMemcpy(iv.data(), R.in_vap, sizeof(int) * dyn.n_in_vap, hipMemcpyDeviceToHost));
  auto fetch = [&](const int *src, int n, int cap, int *dst, int *ndst, int kind) -> int {
    *ndst = n;
    if (!dst || n == 0) return MDN_OK;
    if (n > cap) return set_error(MDN_E_CAPACITY, "mdn_get_state: caller list capacity too small");
    std::vector<int> tmp(2 * (size_t)n);
    hipError_t e = hipMemcpy(tmp.data(), src, sizeof(int) * 2 * n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return set_error(MDN_E_HIP, hipGetErrorString(e));
    for (int k = 0; k < n; k++) {
      dst[2 * k] = ig[tmp[2 * k]];
      dst[2 * k + 1] = kind == 0 ? tmp[2 * k + 1] : (kind == 1 ? iv[tmp[2 * k + 1]] : ig[tmp[2 * k + 1]]);
    }
    return MDN_OK;
  };
  if ((rc = fetch(R.gi, dyn.n_gi, st->cap_gi, st->gi, &st->n_gi, 0))) return rc;
  if ((rc = fetch(R.gv, dyn.n_gv, st->cap_gv, st->gv, &st->n_gv, 1))) return rc;
  if ((rc = fetch(R.gg, dyn.n_gg, st->cap_gg, st->gg, &st->n_gg, 2))) return rc;
  // a pair list that overflowed on the device was truncated: the state is
  // incomplete.  The LDS-capacity flag alone leaves a complete state.
  if (dyn.status & (ST_CAP_GI | ST_CAP_GV | ST_CAP_GG))
    return set_error(MDN_E_CAPACITY, "replica status flags set: a pair list overflowed (lists truncated)");
  return MDN_OK;
}

int mdn_set_positions(mdn_ctx *c, int32_t r, const double *atom_x) {
  if (c) c->mut++;
  int rc = check_replica(c, r);
  if (rc) return rc;
  if (!atom_x) return set_error(MDN_E_ARG, "mdn_set_positions: null array");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(c->reps[r].dev.atom_x, atom_x, sizeof(double) * 3 * (size_t)c->reps[r].def.n_atoms,
                        hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MDN_OK;
}

int mdn_get_forces(mdn_ctx *c, int32_t r, double *atom_f) {
  int rc = check_replica(c, r);
  if (rc) return rc;
  if (!atom_f) return set_error(MDN_E_ARG, "mdn_get_forces: null array");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(atom_f, c->reps[r].dev.atom_f, sizeof(double) * 3 * (size_t)c->reps[r].def.n_atoms,
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MDN_OK;
}

int mdn_force_jvp(mdn_ctx *c, int32_t r, uint32_t bits, const double *atom_u, double *atom_hu) {
  int rc = check_replica(c, r);
  if (rc) return rc;
  if (!atom_u || !atom_hu) return set_error(MDN_E_ARG, "mdn_force_jvp: null array");
  if (bits & (MDN_ION_BMH | MDN_ION_TERSOFF | MDN_ION_SW))
    return set_error(MDN_E_ARG, "mdn_force_jvp: no derivative for the BMH, Tersoff or SW ion terms");
  HIPCHK(hipSetDevice(c->device));
  const size_t n3 = 3 * (size_t)c->reps[r].def.n_atoms;
  if (c->jvp_cap < n3) {
    if (c->d_jvp) HIPCHK(hipFree(c->d_jvp));
    c->d_jvp = nullptr;
    c->jvp_cap = 0;
    HIPCHK(hipMalloc((void **)&c->d_jvp, sizeof(double) * 2 * n3));
    c->jvp_cap = n3;
  }
  double *du = c->d_jvp, *dhu = c->d_jvp + c->jvp_cap;
  HIPCHK(hipMemcpyAsync(du, atom_u, sizeof(double) * n3, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(dhu, 0, sizeof(double) * n3, c->stream));
  hipLaunchKernelGGL(k_force_jvp, dim3(512), dim3(kBlk), 0, c->stream, c->d_reps + r, bits, du, dhu);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(atom_hu, dhu, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MDN_OK;
}

static int refuse_pending(mdn_ctx *c, int32_t r, const char *fn) {
  if (!c->reps[r].pending) return MDN_OK;
  return set_error(MDN_E_STATE, std::string(fn) + ": replica " + std::to_string(r) +
                                    " holds a step left unfinished by mdn_run_batched; finish it with mdn_verlet "
                                    "(or mdn_compute followed by mdn_velocity_calculation)");
}

int mdn_make_pair(mdn_ctx *c, int32_t r) {
  if (c) c->mut++;
  int rc = check_replica(c, r);
  if (rc) return rc;
  if ((rc = refuse_pending(c, r, "mdn_make_pair"))) return rc;
  HIPCHK(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_rebuild, dim3(1), dim3(1024), 0, c->stream, c->d_reps + r, 1);
  HIPCHK(hipGetLastError());
  return MDN_OK;
}

int mdn_zero_forces(mdn_ctx *c, int32_t r) {
  if (c) c->mut++;
  int rc = check_replica(c, r);
  if (rc) return rc;
  const DevReplica &R = c->reps[r].dev;
  HIPCHK(hipMemsetAsync(R.atom_f, 0, sizeof(double) * 3 * (size_t)c->reps[r].def.n_atoms, c->stream));
  HIPCHK(hipMemsetAsync(R.dyn->U, 0, sizeof(double) * U_N, c->stream));
  return MDN_OK;
}

int mdn_compute(mdn_ctx *c, int32_t r, uint32_t bits, int32_t eflag, int32_t precision) {
  if (c) c->mut++;
  int rc = check_replica(c, r);
  if (rc) return rc;
  HIPCHK(hipSetDevice(c->device));
  const HostReplica &h = c->reps[r];
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  rc = precision == MDN_MIXED ? launch_terms<float>(c, h, c->d_reps + r, bits, eflag != 0)
                              : launch_terms<double>(c, h, c->d_reps + r, bits, eflag != 0);
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  c->last_ms = -1;
  return rc;
}

int mdn_velocity_calculation(mdn_ctx *c, int32_t r) {
  if (c) c->mut++;
  int rc = check_replica(c, r);
  if (rc) return rc;
  if (c->reps[r].pending) {  // this kick ends the step left unfinished
    hipLaunchKernelGGL(k_resume, dim3(1), dim3(64), 0, c->stream, c->d_reps + r, 0);
    c->reps[r].pending = 0;
  }
  hipLaunchKernelGGL(k_kick, dim3(grid_for(c->reps[r].def.n_atoms)), dim3(kBlk), 0, c->stream, c->d_reps + r, 0, 0,
                     0);
  HIPCHK(hipGetLastError());
  return MDN_OK;
}

int mdn_update_position(mdn_ctx *c, int32_t r) {
  if (c) c->mut++;
  int rc = check_replica(c, r);
  if (rc) return rc;
  if ((rc = refuse_pending(c, r, "mdn_update_position"))) return rc;
  const HostReplica &h = c->reps[r];
  bool sh = h.def.shake && (h.def.sp[0].rigid.size() || h.def.sp[1].rigid.size() || h.def.sp[2].rigid.size());
  const DevReplica *R = c->d_reps + r;
  hipLaunchKernelGGL(k_drift, dim3(grid_for(h.def.n_atoms)), dim3(kBlk), 0, c->stream, R, sh ? 1 : 0);
  if (sh) launch_shake(c, h, R);
  hipLaunchKernelGGL(k_com_weights, dim3(com_blocks(h.def.n_atoms)), dim3(1024), 0, c->stream, R, 0);
  HIPCHK(hipGetLastError());
  return MDN_OK;
}

int mdn_verlet(mdn_ctx *c, int32_t r, int32_t nsteps, int32_t eflag_every, int64_t step0, int32_t precision) {
  if (c) c->mut++;
  int rc = check_replica(c, r);
  if (rc) return rc;
  HIPCHK(hipSetDevice(c->device));
  HostReplica &h = c->reps[r];
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  for (int s = 0; s < nsteps; s++) {
    bool ef = eflag_every > 0 && ((step0 + s + 1) % eflag_every) == 0;
    if (s == 0 && h.pending)  // the step the batched kernel left unfinished
      rc = precision == MDN_MIXED ? launch_resume<float>(c, h, c->d_reps + r, ef)
                                  : launch_resume<double>(c, h, c->d_reps + r, ef);
    else
      rc = precision == MDN_MIXED ? launch_step<float>(c, h, c->d_reps + r, ef)
                                  : launch_step<double>(c, h, c->d_reps + r, ef);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  c->last_ms = -1;  // read from the events on demand
  return MDN_OK;
}

// mdn_verlet with each step's launch sequence captured once into a hipGraph
// (per replica, precision and energy flag) and replayed
int mdn_verlet_graph(mdn_ctx *c, int32_t r, int32_t nsteps, int32_t eflag_every, int64_t step0, int32_t precision) {
  if (c) c->mut++;
  int rc = check_replica(c, r);
  if (rc) return rc;
  HIPCHK(hipSetDevice(c->device));
  HostReplica &h = c->reps[r];
  const int p = precision == MDN_MIXED ? 1 : 0;
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  for (int s = 0; s < nsteps; s++) {
    const int ef = eflag_every > 0 && ((step0 + s + 1) % eflag_every) == 0;
    if (s == 0 && h.pending) {  // the step the batched kernel left unfinished
      rc = p ? launch_resume<float>(c, h, c->d_reps + r, ef != 0) : launch_resume<double>(c, h, c->d_reps + r, ef != 0);
      if (rc) return rc;
      continue;
    }
    if (!h.graph[p][ef]) {
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      rc = p ? launch_step<float>(c, h, c->d_reps + r, ef != 0) : launch_step<double>(c, h, c->d_reps + r, ef != 0);
      hipError_t e2 = hipStreamEndCapture(c->stream, &g);
      if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
      if (e2 != hipSuccess) return set_error(MDN_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e2));
      hipError_t e3 = hipGraphInstantiate(&h.graph[p][ef], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e3 != hipSuccess) return set_error(MDN_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e3));
    }
    HIPCHK(hipGraphLaunch(h.graph[p][ef], c->stream));
  }
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  c->last_ms = -1;
  return MDN_OK;
}

}  // extern "C"

namespace mdn {
// LDS capacities for a set of replicas from their current in-domain counts,
// with headroom for molecules that enter during a launch
struct BatchCaps { int cap, cap_h, kcap, acap, scap; size_t lds; };
// what a set of replicas needs of a launch's LDS arrays: the largest of each
// replica's counts (in-domain atoms, ion + vapor atoms, their maxima, species
// atoms, exclusion-mask words, constraints)
struct CapNeed {
  int need = 16, need_h = 16, most = 16, most_h = 16, kcap = 8, acap = 1, scap = 1;
  void add(const CapNeed &o) {
    need = std::max(need, o.need); need_h = std::max(need_h, o.need_h);
    most = std::max(most, o.most); most_h = std::max(most_h, o.most_h);
    kcap = std::max(kcap, o.kcap); acap = std::max(acap, o.acap); scap = std::max(scap, o.scap);
  }
};
static CapNeed cap_need(const mdn_ctx *c, const Dyn &dy, int r) {
  const SysDef &d = c->reps[r].def;
  CapNeed q;
  q.scap = std::max(q.scap, (int)std::max(d.sp[0].rigid.size(), d.sp[2].rigid.size()) / 3);
  const int nh = d.sp[0].natoms + dy.n_in_vap * d.sp[2].natoms;
  q.need_h = std::max(q.need_h, nh);
  q.need = std::max(q.need, nh + dy.n_in_gas * d.sp[1].natoms);
  q.most = std::max(q.most, d.n_atoms);
  q.most_h = std::max(q.most_h, d.sp[0].natoms + d.n_vap * d.sp[2].natoms);
  q.kcap = std::max(q.kcap, d.sp[0].natoms + d.sp[1].natoms + d.sp[2].natoms);
  if (d.flags & MDN_ION_AMBER) q.acap = std::max(q.acap, d.sp[0].natoms * ((d.sp[0].natoms + 31) / 32));
  return q;
}
static BatchCaps caps_of(const mdn_ctx *c, const CapNeed &q, size_t rl, size_t budget) {
  // the ion + vapor capacity has a block (64 atoms) of slack past the range
  // (dense_n3 reads whole blocks of positions, weights, forces and
  // parameters), the total a multiple of 32 and at least that
  auto r32 = [](int v) { return (v + 31) / 32 * 32; };
  auto r64 = [](int v) { return (v + 63) / 64 * 64; };
  const int need = q.need, need_h = q.need_h, most = r32(q.most), most_h = r64(q.most_h) + 64;
  int lo_h = std::min(most_h, r64(need_h + 32) + 64), hi_h = std::min(most_h, r64(need_h * 3 / 2 + 32) + 64);
  int lo = std::min(most, r32(need + need / 8 + 64)), hi = std::min(most, r32(need * 3 / 2 + 64));
  BatchCaps b;
  const int kcap = q.kcap + (q.kcap & 1);  // keeps the LDS arrays after it 16-byte aligned
  b.kcap = kcap; b.acap = q.acap; b.scap = q.scap;
  if (c->batch_cap > 0) {
    b.cap = r32(c->batch_cap);
    b.cap_h = std::min(hi_h, b.cap & ~63);
  } else {
    b.cap = hi; b.cap_h = hi_h;
    while (batch_lds_bytes(std::max(b.cap, b.cap_h), b.cap_h, kcap, q.acap, q.scap, rl) > budget &&
           (b.cap > lo || b.cap_h > lo_h)) {
      if (b.cap > lo) b.cap -= 32;
      if (b.cap_h > lo_h) b.cap_h -= 64;
    }
  }
  b.cap = std::max(b.cap, b.cap_h);
  b.lds = batch_lds_bytes(b.cap, b.cap_h, kcap, q.acap, q.scap, rl);
  return b;
}
static BatchCaps batch_caps(const mdn_ctx *c, const Dyn *dyn, const std::vector<int> &set, size_t rl,
                            size_t budget) {
  CapNeed q;
  for (int r : set) q.add(cap_need(c, dyn[r], r));
  return caps_of(c, q, rl, budget);
}

template <typename Rl>
static int launch_batched(mdn_ctx *c, hipStream_t st, const int *order, int n, int threads, int ns, int eflag_every,
                          long long step0, const BatchCaps &b, bool mb) {
  if (n <= 0) return MDN_OK;
  if (b.lds > 160 * 1024)
    return set_error(MDN_E_CAPACITY, "mdn_run_batched: a replica's in-domain atoms exceed one workgroup's LDS; "
                                     "use mdn_verlet for it");
  // launches of different sizes may run concurrently: allow the full LDS
  // (once per kernel variant and context)
  auto kern = mb ? k_batched<Rl, true> : k_batched<Rl, false>;
  const int vbit = 1 << ((sizeof(Rl) == 4 ? 2 : 0) + (mb ? 1 : 0));
  if (!(c->lds_attr & vbit)) {
    HIPCHK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    c->lds_attr |= vbit;
  }
  int grid = c->batch_grid > 0 ? std::min(c->batch_grid, n) : n;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), b.lds, st, c->d_reps, order, n, ns, eflag_every, step0,
                     b.cap, b.cap_h, b.kcap, b.acap, b.scap, c->d_cnt);
  HIPCHK(hipGetLastError());
  return MDN_OK;
}

// a launch over these replicas needs the many-body / BMH variant
static bool needs_manybody(const mdn_ctx *c, const std::vector<int> &set) {
  for (int r : set)
    if (c->reps[r].def.flags & (MDN_ION_BMH | MDN_ION_TERSOFF | MDN_ION_SW)) return true;
  return false;
}
}  // namespace mdn

// XCDs of the device: CDNA3/4 parts are built from 32-CU XCDs (MI355X: 8 x
// 32); other CU counts are taken as one
static int xcds(int n_cu) { return n_cu >= 64 && n_cu % 32 == 0 ? n_cu / 32 : 1; }

// atoms by which a replica's counts may grow before a reused split plan is
// made anew (a rebuild moves a few gas and at most a vapor molecule or two)
constexpr int kPlanSlack = 16;

static void SplitPlan_store(mdn_ctx *c, const std::vector<int> &heavy, const std::vector<int> &clsB,
                            const std::vector<int> &clsA, const std::vector<int> &order, size_t rl) {
  auto &P = c->plan;
  P.valid = true;
  P.age = 0;
  P.mut = c->mut;
  P.rl = rl;
  P.own_threads = c->own_threads;
  P.shared_threads = c->shared_threads;
  P.heavy_factor = c->heavy_factor;
  P.heavy = heavy; P.clsB = clsB; P.clsA = clsA; P.order = order;
  P.need.assign(order.size(), 0);
  P.need_h.assign(order.size(), 0);
}

extern "C" {

int mdn_run_batched(mdn_ctx *c, int32_t nsteps, int32_t eflag_every, int64_t step0, int32_t precision) {
  if (!c) return set_error(MDN_E_ARG, "null context");
  if (!c->committed) return set_error(MDN_E_STATE, "call mdn_commit first");
  HIPCHK(hipSetDevice(c->device));
  for (auto &h : c->reps)
    if (h.def.shake && h.def.sp[1].rigid.size())
      return set_error(MDN_E_ARG, "mdn_run_batched: constrained gas molecules are not supported; use mdn_verlet");
  const size_t rl = precision == MDN_MIXED ? sizeof(float) : sizeof(double);
  const int n = (int)c->reps.size();
  const int chunk = c->batch_chunk > 0 ? c->batch_chunk : nsteps;
  if (c->h_cap < n) {
    if (c->h_dyn) HIPCHK(hipHostFree(c->h_dyn));
    if (c->h_order) HIPCHK(hipHostFree(c->h_order));
    if (c->h_cnt) HIPCHK(hipHostFree(c->h_cnt));
    c->h_dyn = nullptr; c->h_order = nullptr; c->h_cnt = nullptr; c->d_cnt = nullptr; c->h_cap = 0;
    HIPCHK(hipHostMalloc((void **)&c->h_dyn, sizeof(Dyn) * n, hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&c->h_order, sizeof(int) * n, hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&c->h_cnt, sizeof(int4) * n, hipHostMallocDefault));
    HIPCHK(hipHostGetDevicePointer((void **)&c->d_cnt, c->h_cnt, 0));
    c->h_cap = n;
  }
  // the counts of every replica at the start of each chunk: read back
  // behind the previous chunk's launches on the same stream
  // h_dyn: the device's Dyn records after a read-back; the batched launches
  // then keep their counts and pending flags current (from h_cnt)
  const Dyn *dyn = c->h_dyn;
  std::vector<int> order(n);
  c->last_ms = 0;
  if (c->dyn_mut != c->mut) {
    HIPCHK(hipMemcpyAsync(c->h_dyn, c->d_dyn, sizeof(Dyn) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  c->dyn_mut = ~0ull;
  for (int done = 0; done < nsteps; done += chunk) {
    const int ns = std::min(chunk, nsteps - done);
    // Work per step grows with the square of the ion + vapor atom count.  The
    // heaviest replicas get a whole CU each (512-thread workgroups: at 256
    // VGPRs one fills a CU); the rest share CUs in 128-thread workgroups, in
    // two classes whose LDS lets four (<= 40 KB) or three (<= 53 KB) share a
    // CU.  The heavy count is the largest (among replicas well above the
    // median cost) that lets every workgroup be resident at once; the CUs
    // left over then go to the costliest of the rest, which move from four
    // to three per CU while the CU count allows.  Workgroups of a launch are
    // dealt round-robin over the XCDs (32 CUs each), so the count that must
    // fit is that of the fullest XCD.  Within a launch the replicas go in
    // descending cost: a workgroup that finds no room starts when one ends,
    // and the cheapest can best afford that.
    const size_t kA = 40 * 1024, kB = 53 * 1024;
    std::vector<int> heavy, clsA, clsB;
    const bool split = c->batch_split && c->batch_threads == 256 && n > c->n_cu;
    auto needs = [&](int r, int &need, int &need_h) {
      const SysDef &d = c->reps[r].def;
      need_h = d.sp[0].natoms + dyn[r].n_in_vap * d.sp[2].natoms;
      need = need_h + dyn[r].n_in_gas * d.sp[1].natoms;
    };
    bool reuse = split && c->plan.valid && c->plan.age + ns <= chunk && c->plan.mut == c->mut && c->plan.rl == rl &&
                 c->plan.own_threads == c->own_threads && c->plan.shared_threads == c->shared_threads &&
                 c->plan.heavy_factor == c->heavy_factor && (int)c->plan.order.size() == n;
    for (int r = 0; r < n && reuse; r++) {
      int need, need_h;
      needs(r, need, need_h);
      reuse = need <= c->plan.need[r] + kPlanSlack && need_h <= c->plan.need_h[r] + kPlanSlack;
    }
    std::vector<double> cost(n);
    if (!reuse) {
      for (int r = 0; r < n; r++) {
        const SysDef &d = c->reps[r].def;
        double nh = d.sp[0].natoms + dyn[r].n_in_vap * d.sp[2].natoms;
        double nin = nh + dyn[r].n_in_gas * d.sp[1].natoms;
        cost[r] = nh * nh + 40.0 * nin;
        order[r] = r;
      }
      std::stable_sort(order.begin(), order.end(), [&](int a, int b2) { return cost[a] > cost[b2]; });
    }
    if (reuse) {
      heavy = c->plan.heavy;
      clsB = c->plan.clsB;
      clsA = c->plan.clsA;
      order = c->plan.order;
    } else if (split) {
      // each replica's own LDS need at the smallest headroom
      std::vector<CapNeed> rq(n);
      std::vector<size_t> own(n);
      for (int r = 0; r < n; r++) {
        rq[r] = cap_need(c, dyn[r], r);
        own[r] = caps_of(c, rq[r], rl, 0).lds;
      }
      double med = cost[order[n / 2]];
      int Hmax = 0;
      while (Hmax < n && Hmax < c->n_cu && cost[order[Hmax]] > c->heavy_factor * med) Hmax++;
      const int nx = xcds(c->n_cu), per = c->n_cu / nx;
      auto up = [&](int k) { return (k + nx - 1) / nx; };
      // workgroups per CU of a class: its LDS (the largest need of its
      // replicas, at the class budget), 8 waves per CU at 256 VGPRs, at most
      // four
      const int wmax = std::min(4, 8 / (c->shared_threads / 64));
      auto per_cu = [&](int cnt, const CapNeed &q, size_t budget) {
        if (cnt == 0) return wmax;
        const size_t l = (caps_of(c, q, rl, budget).lds + 1023) / 1024 * 1024;
        return (int)std::min<size_t>(wmax, (160 * 1024) / l);
      };
      // CUs of the fullest XCD
      auto cus = [&](int nh, int na, const CapNeed &qa, int nb, const CapNeed &qb) {
        const int fa = per_cu(na, qa, kA), fb = per_cu(nb, qb, kB);
        if (fa == 0 || fb == 0) return per + 1;
        return up(nh) + (up(na) + fa - 1) / fa + (up(nb) + fb - 1) / fb;
      };
      // the classes for H replicas with a CU of their own; true when every
      // workgroup can be resident at once.  Fewer own-CU replicas need
      // fewer CUs, so the largest H that fits is found by bisection; when
      // not even H = 0 fits, its classes are used as they are.
      std::vector<CapNeed> sufA;
      CapNeed qb;
      auto attempt = [&](int H) {
        heavy.assign(order.begin(), order.begin() + H);
        clsA.clear();
        clsB.clear();
        for (int k = H; k < n; k++) {
          const int r = order[k];
          if (own[r] <= kA) clsA.push_back(r);
          else if (own[r] <= kB) clsB.push_back(r);
          else heavy.push_back(r);
        }
        // needs of clsA's suffixes (its costliest leave from the front)
        const int na = (int)clsA.size();
        sufA.assign(na + 1, CapNeed());
        for (int k = na - 1; k >= 0; k--) { sufA[k] = sufA[k + 1]; sufA[k].add(rq[clsA[k]]); }
        qb = CapNeed();
        for (int r : clsB) qb.add(rq[r]);
        return cus((int)heavy.size(), na, sufA[0], (int)clsB.size(), qb) <= per;
      };
      if (attempt(0)) {
        int lo = 0, hi = Hmax;
        while (lo < hi) {
          const int mid = (lo + hi + 1) / 2;
          if (attempt(mid)) lo = mid;
          else hi = mid - 1;
        }
        attempt(lo);
        // move clsA's costliest (it is in descending cost order) to clsB
        // while the fullest XCD still holds every workgroup
        const int nh = (int)heavy.size(), na = (int)clsA.size(), nb = (int)clsB.size();
        int moved = 0;
        while (moved < na) {
          CapNeed q2 = qb;
          q2.add(rq[clsA[moved]]);
          if (cus(nh, na - moved - 1, sufA[moved + 1], nb + moved + 1, q2) > per) break;
          qb = q2;
          moved++;
        }
        clsB.insert(clsB.end(), clsA.begin(), clsA.begin() + moved);
        clsA.erase(clsA.begin(), clsA.begin() + moved);
      }
      auto by_cost = [&](int a2, int b2) { return cost[a2] > cost[b2]; };
      std::stable_sort(heavy.begin(), heavy.end(), by_cost);
      std::stable_sort(clsB.begin(), clsB.end(), by_cost);
      order.clear();
      order.insert(order.end(), heavy.begin(), heavy.end());
      order.insert(order.end(), clsB.begin(), clsB.end());
      order.insert(order.end(), clsA.begin(), clsA.end());
      SplitPlan_store(c, heavy, clsB, clsA, order, rl);
      for (int r = 0; r < n; r++) needs(r, c->plan.need[r], c->plan.need_h[r]);
    }
    if (split) c->plan.age += ns;
    const int H = (int)heavy.size(), NB_ = (int)clsB.size(), NA = (int)clsA.size();
    // the launch order on the device (still there when the plan is reused)
    const bool upload = !(reuse && c->plan_on_device);
    if (upload) {
      memcpy(c->h_order, order.data(), sizeof(int) * n);
      HIPCHK(hipMemcpyAsync(c->d_order, c->h_order, sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
    }
    c->plan_on_device = split;
    const long long s0 = (long long)(step0 + done);
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    int rc;
    if (H + NB_ + NA > 0) {
      // heavy on stream, class B on stream2, class A on stream3, all after
      // an order upload of this chunk.  Otherwise the streams hold nothing
      // else: the previous call ended with all three synchronized, and a
      // state change since then made this call read back (and synchronize)
      // first
      if (upload) {
        HIPCHK(hipEventRecord(c->evj, c->stream));
        HIPCHK(hipStreamWaitEvent(c->stream2, c->evj, 0));
        HIPCHK(hipStreamWaitEvent(c->stream3, c->evj, 0));
      }
      size_t lds = 0;
      int cap = 0;
      struct Part { hipStream_t st; int off, cnt, threads; size_t budget; const std::vector<int> *set; };
      const Part parts[3] = {{c->stream, 0, H, c->own_threads, 160 * 1024, &heavy},
                             {c->stream2, H, NB_, c->shared_threads, kB, &clsB},
                             {c->stream3, H + NB_, NA, c->shared_threads, kA, &clsA}};
      for (const Part &pt : parts) {
        if (pt.cnt == 0) continue;
        BatchCaps bc = batch_caps(c, dyn, *pt.set, rl, pt.budget);
        rc = precision == MDN_MIXED
                 ? launch_batched<float>(c, pt.st, c->d_order + pt.off, pt.cnt, pt.threads, ns, eflag_every, s0, bc,
                                         needs_manybody(c, *pt.set))
                 : launch_batched<double>(c, pt.st, c->d_order + pt.off, pt.cnt, pt.threads, ns, eflag_every, s0, bc,
                                          needs_manybody(c, *pt.set));
        if (rc) return rc;
        lds = std::max(lds, bc.lds);
        cap = std::max(cap, bc.cap);
      }
      HIPCHK(hipEventRecord(c->evk, c->stream2));
      HIPCHK(hipEventRecord(c->evl, c->stream3));
      c->last_cap = cap;
      c->last_lds = (int)lds;
    } else {
      BatchCaps b = batch_caps(c, dyn, order, rl, 80 * 1024);
      rc = precision == MDN_MIXED
               ? launch_batched<float>(c, c->stream, c->d_order, n, c->batch_threads, ns, eflag_every, s0, b,
                                       needs_manybody(c, order))
               : launch_batched<double>(c, c->stream, c->d_order, n, c->batch_threads, ns, eflag_every, s0, b,
                                        needs_manybody(c, order));
      if (rc) return rc;
      c->last_cap = b.cap;
      c->last_lds = (int)b.lds;
    }
    c->last_heavy = H;
    c->last_classes[0] = H;
    c->last_classes[1] = NB_;
    c->last_classes[2] = NA;
    c->last_classes[3] = H + NB_ + NA > 0 ? 0 : n;
    c->last_order = order;
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    // the chunk's end: each stream synchronized (no joins queued behind the
    // kernels); the kernel time runs from ev0 to the last of the three ends
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0, m2 = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (H + NB_ + NA > 0) {
      HIPCHK(hipStreamSynchronize(c->stream2));
      HIPCHK(hipStreamSynchronize(c->stream3));
      HIPCHK(hipEventElapsedTime(&m2, c->ev0, c->evk));
      ms = std::max(ms, m2);
      HIPCHK(hipEventElapsedTime(&m2, c->ev0, c->evl));
      ms = std::max(ms, m2);
    }
    c->last_ms += ms;
    for (int r = 0; r < n; r++) {
      const int4 q = c->h_cnt[r];
      c->h_dyn[r].n_in_gas = q.x;
      c->h_dyn[r].n_in_vap = q.y;
      c->h_dyn[r].pending = q.z;
    }
  }
  // replicas stopped inside a step (LDS capacity outgrown at a rebuild): the
  // per-call path finishes that step first
  for (int r = 0; r < n; r++) c->reps[r].pending = c->h_dyn[r].pending;
  c->dyn_mut = c->mut;
  return MDN_OK;
}

int mdn_set_batch_shape(mdn_ctx *c, int32_t threads, int32_t lds_cap_atoms, int32_t grid, int32_t chunk) {
  if (!c || (threads != 0 && threads != 64 && threads != 128 && threads != 256 && threads != 512))
    return set_error(MDN_E_ARG, "mdn_set_batch_shape: threads must be 0 (split), 64, 128, 256 or 512");
  c->batch_split = threads == 0 ? 1 : 0;
  c->batch_threads = threads == 0 ? 256 : threads;
  c->batch_cap = lds_cap_atoms;
  c->batch_grid = grid;
  c->batch_chunk = chunk;
  return MDN_OK;
}

int mdn_set_batch_tuning(mdn_ctx *c, int32_t own_threads, int32_t shared_threads, double heavy_factor) {
  if (!c) return set_error(MDN_E_ARG, "null context");
  if (own_threads != 256 && own_threads != 512)
    return set_error(MDN_E_ARG, "mdn_set_batch_tuning: own_threads must be 256 or 512");
  if (shared_threads != 64 && shared_threads != 128 && shared_threads != 256)
    return set_error(MDN_E_ARG, "mdn_set_batch_tuning: shared_threads must be 64, 128 or 256");
  if (!(heavy_factor >= 1.0 && heavy_factor <= 100.0))
    return set_error(MDN_E_ARG, "mdn_set_batch_tuning: heavy_factor must lie in [1, 100]");
  c->own_threads = own_threads;
  c->shared_threads = shared_threads;
  c->heavy_factor = heavy_factor;
  return MDN_OK;
}

int mdn_get_counters(mdn_ctx *c, int32_t r, int64_t *atom_steps, int64_t *rebuilds, int32_t *shake_iters_max,
                     int64_t *shake_failures) {
  int rc = check_replica(c, r);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  Dyn dyn;
  HIPCHK(hipMemcpy(&dyn, c->reps[r].dev.dyn, sizeof(Dyn), hipMemcpyDeviceToHost));
  if (atom_steps) *atom_steps = dyn.atom_steps;
  if (rebuilds) *rebuilds = dyn.rebuild_count;
  if (shake_iters_max) *shake_iters_max = dyn.shake_iters_max;
  if (shake_failures) *shake_failures = (int64_t)dyn.shake_fail;
  int code = status_to_code(dyn.status);
  return code ? set_error(code, "replica status flags set: " + std::to_string(dyn.status)) : MDN_OK;
}

int mdn_get_batch_info(mdn_ctx *c, int32_t *lds_cap_atoms, int32_t *lds_bytes, int32_t *heavy) {
  if (!c) return set_error(MDN_E_ARG, "null context");
  if (heavy) *heavy = c->last_heavy;
  if (lds_cap_atoms) *lds_cap_atoms = c->last_cap;
  if (lds_bytes) *lds_bytes = c->last_lds;
  return MDN_OK;
}

int mdn_get_batch_plan(mdn_ctx *c, int32_t *counts, int32_t *order, int32_t cap) {
  if (!c || !counts) return set_error(MDN_E_ARG, "null argument");
  for (int k = 0; k < 4; k++) counts[k] = c->last_classes[k];
  if (order) {
    if (cap < (int)c->last_order.size())
      return set_error(MDN_E_CAPACITY, "mdn_get_batch_plan: order holds fewer entries than replicas");
    for (size_t k = 0; k < c->last_order.size(); k++) order[k] = c->last_order[k];
  }
  return MDN_OK;
}

int mdn_sync(mdn_ctx *c) {
  if (!c) return set_error(MDN_E_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  return MDN_OK;
}

int mdn_last_kernel_ms(mdn_ctx *c, double *ms) {
  if (!c || !ms) return set_error(MDN_E_ARG, "null argument");
  if (c->last_ms >= 0) {
    *ms = c->last_ms;
    return MDN_OK;
  }
  float f = 0;
  HIPCHK(hipEventSynchronize(c->ev1));
  HIPCHK(hipEventElapsedTime(&f, c->ev0, c->ev1));
  *ms = f;
  return MDN_OK;
}

}  // extern "C"

// ======================= trajectory-segment gradient =====================
// Reverse-mode gradient of a trajectory loss over one segment of MD::verlet
// (verlet.cpp:20-39 as shipped: half kick, drift, forces, half kick; no
// SHAKE, no rebuild within the segment, blend weights 1) with respect to the
// ion's per-atom sigma, eps, q of the vi term (VaporIon.cpp:5-43), every
// array in device memory.  The segment's atoms are the ion atoms and every
// vapor atom (molecule order); their positions reach the replica's atom
// arrays for the per-call force kernels of the other terms, and the vi term
// runs over the segment's own arrays with the parameters as inputs.
//
// One step n -> n + 1:
//   vh = v_n + (c F(x_n)) / m,  x_{n+1} = x_n + dt vh,  v_{n+1} = vh + (c F(x_{n+1})) / m
// with c = 0.5 dt 4.184e-4 (verlet.cpp:49).  With a_x, a_v the adjoints of
// x_{n+1}, v_{n+1}, one step back is
//   a_v += dt a_x;  u = a_v / m;  a_x += k c H(x_n) u (+ dL/dx_n);
//   g += k c (dF(x_n)/dtheta)^T u
// (k = 2 for n > 0: x_n enters two half kicks with the same u; k = 1 for
// n = 0).  H(x) u is the derivative of the forces along u: the vi term's
// closed form and the dual-number instantiation of the per-call force
// kernels (k_force_jvp).  The forward pass keeps (x, v) at the start of each
// checkpoint interval; the reverse pass recomputes an interval's positions
// from its checkpoint, runs its steps back, and takes the interval's
// parameter gradient in one adjoint call over its frames.
namespace mdn {

constexpr int kSegBlk = 256;

static int seg_grid(size_t n) { return (int)std::max<size_t>(1, (n + kSegBlk - 1) / kSegBlk); }

// dst[3 idx[a] + q] = src[3 a + q]
__global__ void k_seg_scatter(const int *__restrict__ idx, int n, const double *__restrict__ src,
                              double *__restrict__ dst) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * n) return;
  const int a = t / 3, q = t - 3 * a;
  dst[3 * (size_t)idx[a] + q] = src[t];
}

// dst[3 a + q] = src[3 idx[a] + q], or += with add
__global__ void k_seg_gather(const int *__restrict__ idx, int n, const double *__restrict__ src,
                             double *__restrict__ dst, int add) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * n) return;
  const int a = t / 3, q = t - 3 * a;
  const double s = src[3 * (size_t)idx[a] + q];
  dst[t] = add ? dst[t] + s : s;
}

// half kick: out = v + (c F) / m
__global__ void k_seg_kick(int n, const double *__restrict__ v, const double *__restrict__ F,
                           const double *__restrict__ m, double c, double *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * n) return;
  out[t] = v[t] + c * F[t] / m[t / 3];
}

// a step's first half: vh = v + (c F) / m and xn = x + dt vh, xn scattered
// into the replica's positions ax (global atom order), and the replica's
// forces af (n_af values) cleared for the force kernels that follow (one
// launch instead of four)
__global__ void k_seg_kick_drift(int n, const double *__restrict__ v, const double *__restrict__ F,
                                 const double *__restrict__ m, double c, const double *__restrict__ x, double dt,
                                 double *__restrict__ vh, double *__restrict__ xn, const int *__restrict__ idx,
                                 double *__restrict__ ax, double *__restrict__ af, int n_af) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 3 * n) {
    const int a = t / 3, q = t - 3 * a;
    const double u = v[t] + c * F[t] / m[a];
    const double y = x[t] + dt * u;
    vh[t] = u;
    xn[t] = y;
    ax[3 * (size_t)idx[a] + q] = y;
  }
  if (t < n_af) af[t] = 0.0;
}

// a step's second half: F += the replica's forces af (gathered from global
// atom order), then vn = vh + (c F) / m
__global__ void k_seg_gather_kick(int n, const int *__restrict__ idx, const double *__restrict__ af,
                                  double *__restrict__ F, const double *__restrict__ vh,
                                  const double *__restrict__ m, double c, double *__restrict__ vn) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * n) return;
  const int a = t / 3, q = t - 3 * a;
  const double f = F[t] + af[3 * (size_t)idx[a] + q];
  F[t] = f;
  vn[t] = vh[t] + c * f / m[a];
}

// an HVP's set-up: x into the replica's positions, u into the direction
// array gu (both global atom order), the result array ghu (n_h values)
// cleared
__global__ void k_seg_hvp_setup(const int *__restrict__ idx, int n, const double *__restrict__ x,
                                double *__restrict__ ax, const double *__restrict__ u, double *__restrict__ gu,
                                double *__restrict__ ghu, int n_h) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n_h) ghu[t] = 0.0;
  if (t < 3 * n) {
    const int a = t / 3, q = t - 3 * a;
    const size_t g = 3 * (size_t)idx[a] + q;
    ax[g] = x[t];
    gu[g] = u[t];
  }
}

// out[0] = sum of (x - y)^2 over n3 values, in a fixed order (one workgroup)
__global__ void __launch_bounds__(kSegBlk) k_seg_sqdist(int n3, const double *__restrict__ x,
                                                        const double *__restrict__ y, double *__restrict__ out) {
  __shared__ double s[kSegBlk];
  double acc = 0;
  for (int t = threadIdx.x; t < n3; t += kSegBlk) {
    const double d = x[t] - y[t];
    acc += d * d;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kSegBlk / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

// a step back, first half: a_v += dt a_x; u = a_v / m; us = kc u
__global__ void k_seg_back_u(int n, double *__restrict__ a_v, const double *__restrict__ a_x,
                             const double *__restrict__ m, double dt, double kc, double *__restrict__ u,
                             double *__restrict__ us) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * n) return;
  const double av = a_v[t] + dt * a_x[t];
  a_v[t] = av;
  const double uu = av / m[t / 3];
  u[t] = uu;
  us[t] = kc * uu;
}

// a step back, second half: a_x += kc H u, then + 2 (x - target) / N when
// the step is observed (target != null)
__global__ void k_seg_back_x(int n3, double *__restrict__ a_x, const double *__restrict__ hu, double kc,
                             const double *__restrict__ x, const double *__restrict__ tgt, double N) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n3) return;
  double a = a_x[t] + kc * hu[t];
  if (tgt) a = a + 2.0 * (x[t] - tgt[t]) / N;
  a_x[t] = a;
}

// a_x = 2 (x - target) / N
__global__ void k_seg_dloss(int n3, const double *__restrict__ x, const double *__restrict__ tgt, double N,
                            double *__restrict__ a_x) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n3) a_x[t] = 2.0 * (x[t] - tgt[t]) / N;
}

__global__ void k_seg_add(int n, double *__restrict__ g, const double *__restrict__ d) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) g[t] += d[t];
}

__global__ void k_seg_fill(size_t n, double *__restrict__ p, double v) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

// device buffers and launch helpers of one mdn_segment_grad call
struct Segment {
  mdn_ctx *c = nullptr;
  const HostReplica *h = nullptr;
  const DevReplica *Rd = nullptr;  // device copy of the replica's descriptor
  DevReplica R;                    // host copy (its device pointers)
  int n_ion = 0, n_vap = 0, N = 0, n_atoms = 0;
  uint32_t rest = 0;  // terms other than vi (mdn_compute bits)
  bool vi = false;
  char *buf = nullptr;
  int *idx = nullptr;
  double *mass = nullptr, *ones = nullptr, *theta = nullptr, *vap = nullptr, *epart = nullptr;
  double *gU = nullptr, *gHU = nullptr;  // direction and result in the replica's global atom order
  ~Segment() {
    if (buf) (void)hipFree(buf);
  }
  ViBatch batch(const double *x, int S) const {
    ViBatch b;
    b.x = x; b.w = ones; b.S = S; b.n_ion = n_ion; b.n_vap = n_vap;
    b.sig_i = theta; b.eps_i = theta + n_ion; b.q_i = theta + 2 * n_ion;
    b.sig_v = vap; b.eps_v = vap + n_vap; b.q_v = vap + 2 * n_vap;
    b.vin = nullptr;
    b.K = mdn_qqrd2e();
    return b;
  }
  // F = forces of the selected terms at the segment positions x
  int forces(const double *x, double *F) {
    const int g = seg_grid(3 * (size_t)N);
    hipLaunchKernelGGL(k_seg_scatter, dim3(g), dim3(kSegBlk), 0, c->stream, idx, N, x, R.atom_x);
    if (rest) HIPCHK(hipMemsetAsync(R.atom_f, 0, sizeof(double) * 3 * (size_t)n_atoms, c->stream));
    return forces_staged(x, F, true);
  }
  // the same once x is in the replica's positions and its forces are clear;
  // without gather, the replica's forces are left for the caller to add
  int forces_staged(const double *x, double *F, bool gather) {
    const int g = seg_grid(3 * (size_t)N);
    if (rest) {
      int rc = launch_terms<double>(c, *h, Rd, rest, false);
      if (rc) return rc;
    }
    if (vi) {
      int rc = vi_forward_dev(batch(x, 1), F, epart, nullptr, c->stream);
      if (rc) return rc;
    } else {
      HIPCHK(hipMemsetAsync(F, 0, sizeof(double) * 3 * (size_t)N, c->stream));
    }
    if (rest && gather) hipLaunchKernelGGL(k_seg_gather, dim3(g), dim3(kSegBlk), 0, c->stream, idx, N, R.atom_f, F, 1);
    HIPCHK(hipGetLastError());
    return MDN_OK;
  }
  // hu = H(x) u, the derivative of the forces along u
  int hvp(const double *x, const double *u, double *hu) {
    const int g = seg_grid(3 * (size_t)N);
    if (vi) {
      int rc = vi_jvp_dev(batch(x, 1), u, hu, c->stream);
      if (rc) return rc;
    } else {
      HIPCHK(hipMemsetAsync(hu, 0, sizeof(double) * 3 * (size_t)N, c->stream));
    }
    const uint32_t bits = rest & ~(uint32_t)MDN_EFIELD;  // the Efield force does not depend on x
    if (bits) {
      const int n_h = 3 * n_atoms, gs = seg_grid(std::max<size_t>(3 * (size_t)N, (size_t)n_h));
      hipLaunchKernelGGL(k_seg_hvp_setup, dim3(gs), dim3(kSegBlk), 0, c->stream, idx, N, x, R.atom_x, u, gU, gHU,
                         n_h);
      hipLaunchKernelGGL(k_force_jvp, dim3(512), dim3(kBlk), 0, c->stream, Rd, bits, gU, gHU);
      hipLaunchKernelGGL(k_seg_gather, dim3(g), dim3(kSegBlk), 0, c->stream, idx, N, gHU, hu, 1);
    }
    HIPCHK(hipGetLastError());
    return MDN_OK;
  }
  // one MD::verlet step: (x, v, F) -> (xn, vn, F); vh is scratch
  int step(const double *x, const double *v, double *F, double *vh, double *xn, double *vn, double cc, double dt) {
    const int g = seg_grid(3 * (size_t)N);
    const int n_af = rest ? 3 * n_atoms : 0, gk = seg_grid(std::max<size_t>(3 * (size_t)N, (size_t)n_af));
    hipLaunchKernelGGL(k_seg_kick_drift, dim3(gk), dim3(kSegBlk), 0, c->stream, N, v, F, mass, cc, x, dt, vh, xn, idx,
                       R.atom_x, R.atom_f, n_af);
    int rc = forces_staged(xn, F, false);
    if (rc) return rc;
    if (rest)
      hipLaunchKernelGGL(k_seg_gather_kick, dim3(g), dim3(kSegBlk), 0, c->stream, N, idx, R.atom_f, F, vh, mass, cc,
                         vn);
    else
      hipLaunchKernelGGL(k_seg_kick, dim3(g), dim3(kSegBlk), 0, c->stream, N, vh, F, mass, cc, vn);
    HIPCHK(hipGetLastError());
    return MDN_OK;
  }
};

}  // namespace mdn

extern "C" {

int mdn_segment_grad(mdn_ctx *c, int32_t r, int32_t nsteps, int32_t every, const double *x0, const double *v0,
                     const double *ion_sigma, const double *ion_eps, const double *ion_q, const int32_t *obs,
                     int32_t n_obs, const double *targets, double *loss, double *g_sigma, double *g_eps,
                     double *g_q, double *x_end, double *v_end) {
  if (c) c->mut++;
  int rc = check_replica(c, r);
  if (rc) return rc;
  if ((rc = refuse_pending(c, r, "mdn_segment_grad"))) return rc;
  if (nsteps < 0 || every < 1) return set_error(MDN_E_ARG, "mdn_segment_grad: nsteps >= 0 and every >= 1 required");
  if (!ion_sigma || !ion_eps || !ion_q || !loss) return set_error(MDN_E_ARG, "mdn_segment_grad: null argument");
  if (n_obs < 0 || (n_obs && (!obs || !targets)))
    return set_error(MDN_E_ARG, "mdn_segment_grad: observed steps need obs and targets");
  const bool grad = g_sigma || g_eps || g_q;
  if (grad && !(g_sigma && g_eps && g_q))
    return set_error(MDN_E_ARG, "mdn_segment_grad: give all three gradient outputs or none");
  std::vector<int> tgt_of((size_t)nsteps + 1, -1);
  for (int k = 0; k < n_obs; k++) {
    if (obs[k] < 0 || obs[k] > nsteps || (k && obs[k] <= obs[k - 1]))
      return set_error(MDN_E_ARG, "mdn_segment_grad: obs must ascend strictly within [0, nsteps]");
    tgt_of[obs[k]] = k;
  }
  HIPCHK(hipSetDevice(c->device));
  const HostReplica &h = c->reps[r];
  const SysDef &d = h.def;
  if (d.shake) return set_error(MDN_E_ARG, "mdn_segment_grad: the segment follows MD::verlet without SHAKE");
  const uint32_t nodiff = MDN_ION_BMH | MDN_ION_TERSOFF | MDN_ION_SW;
  if (grad && (d.flags & nodiff))
    return set_error(MDN_E_ARG, "mdn_segment_grad: no derivative for the BMH, Tersoff or SW ion terms");
  Segment sg;
  sg.c = c; sg.h = &h; sg.Rd = c->d_reps + r; sg.R = h.dev;
  sg.n_ion = d.sp[0].natoms;
  sg.n_vap = d.n_vap * d.sp[2].natoms;
  sg.N = sg.n_ion + sg.n_vap;
  sg.n_atoms = d.n_atoms;
  sg.vi = (d.flags & MDN_VI) != 0;
  sg.rest = d.flags & (MDN_ION_AMBER | MDN_ION_BMH | MDN_ION_TERSOFF | MDN_ION_SW | MDN_VV | MDN_VAPOR_INTRA |
                       MDN_EFIELD);
  const int N = sg.N, N3 = 3 * N, ni = sg.n_ion, nv = sg.n_vap;
  const int g0 = d.sp[0].natoms + d.n_gas * d.sp[1].natoms;  // first vapor atom
  std::vector<int> idx(N);
  for (int a = 0; a < ni; a++) idx[a] = a;
  for (int a = 0; a < nv; a++) idx[ni + a] = g0 + a;
  // the segment's conditions on the replica's current state: every vapor
  // molecule in the domain, blend weight 1 on every segment atom
  {
    std::vector<uint8_t> reg(d.n_mol);
    std::vector<double> w(d.n_atoms);
    HIPCHK(hipMemcpyAsync(reg.data(), h.dev.region, d.n_mol, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(w.data(), h.dev.atom_w, sizeof(double) * d.n_atoms, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int m = 1 + d.n_gas; m < d.n_mol; m++)
      if (!reg[m]) return set_error(MDN_E_ARG, "mdn_segment_grad: every vapor molecule must be in the domain");
    for (int a = 0; a < N; a++)
      if (w[idx[a]] != 1.0)
        return set_error(MDN_E_ARG, "mdn_segment_grad: every segment atom must have blend weight 1");
  }
  const int nint = (nsteps + every - 1) / every;
  const int S = std::max(1, std::min(every, std::max(nsteps, 1)));  // frames per interval
  // one allocation, carved twice: first to size it, then for real
  double *xa, *xb, *vv, *vh, *F, *u, *hu, *ax, *av, *ck, *frames, *us, *tg, *gsum, *gint, *lp, *save_x, *save_f;
  size_t used = 0;
  auto carve = [&](double *base) {
    used = 0;
    auto take = [&](size_t n) { double *q = base ? base + used : nullptr; used += n; return q; };
    xa = take(N3); xb = take(N3); vv = take(N3); vh = take(N3); F = take(N3);
    u = take(N3); hu = take(N3); ax = take(N3); av = take(N3);
    ck = take((size_t)2 * nint * N3);
    frames = take((size_t)S * N3); us = take((size_t)S * N3);
    tg = take((size_t)n_obs * N3);
    sg.mass = take(N);
    sg.ones = take((size_t)S * N);
    sg.theta = take(3 * (size_t)ni);
    sg.vap = take(3 * (size_t)nv);
    sg.epart = take(3 * (size_t)S * ni);
    gsum = take(3 * (size_t)ni); gint = take(3 * (size_t)ni);
    lp = take((size_t)nsteps + 1);
    sg.gU = take(3 * (size_t)d.n_atoms); sg.gHU = take(3 * (size_t)d.n_atoms);
    save_x = take(3 * (size_t)d.n_atoms); save_f = take(3 * (size_t)d.n_atoms);
  };
  carve(nullptr);
  const size_t nd = used;
  HIPCHK(hipMalloc((void **)&sg.buf, sizeof(double) * nd + sizeof(int) * (size_t)N));
  carve(reinterpret_cast<double *>(sg.buf));
  sg.idx = reinterpret_cast<int *>(sg.buf + sizeof(double) * nd);
  // host tables: masses, the vapor's per-atom parameters (type sigma, eps;
  // charge), the ion parameters
  std::vector<double> hm(N), hv(3 * (size_t)nv), ht(3 * (size_t)ni);
  for (int a = 0; a < ni; a++) hm[a] = d.sp[0].mass[a];
  const Species &V = d.sp[2];
  for (int a = 0; a < nv; a++) {
    const int k = a % V.natoms;
    hm[ni + a] = V.mass[k];
    hv[a] = d.t_sig[V.type[k]];
    hv[nv + a] = d.t_eps[V.type[k]];
    hv[2 * nv + a] = V.charge[k];
  }
  for (int a = 0; a < ni; a++) { ht[a] = ion_sigma[a]; ht[ni + a] = ion_eps[a]; ht[2 * ni + a] = ion_q[a]; }
  hipStream_t st = c->stream;
  HIPCHK(hipMemcpyAsync(sg.idx, idx.data(), sizeof(int) * N, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(sg.mass, hm.data(), sizeof(double) * N, hipMemcpyHostToDevice, st));
  if (nv) HIPCHK(hipMemcpyAsync(sg.vap, hv.data(), sizeof(double) * 3 * nv, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(sg.theta, ht.data(), sizeof(double) * 3 * ni, hipMemcpyHostToDevice, st));
  if (n_obs) HIPCHK(hipMemcpyAsync(tg, targets, sizeof(double) * N3 * (size_t)n_obs, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_seg_fill, dim3(seg_grid((size_t)S * N)), dim3(kSegBlk), 0, st, (size_t)S * N, sg.ones, 1.0);
  hipLaunchKernelGGL(k_seg_fill, dim3(seg_grid(3 * (size_t)d.n_atoms)), dim3(kSegBlk), 0, st,
                     3 * (size_t)d.n_atoms, sg.gU, 0.0);
  // the replica's positions and forces are restored at the end
  HIPCHK(hipMemcpyAsync(save_x, h.dev.atom_x, sizeof(double) * 3 * d.n_atoms, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(save_f, h.dev.atom_f, sizeof(double) * 3 * d.n_atoms, hipMemcpyDeviceToDevice, st));
  const int gN = seg_grid(3 * (size_t)N);
  if (x0) HIPCHK(hipMemcpyAsync(xa, x0, sizeof(double) * N3, hipMemcpyHostToDevice, st));
  else hipLaunchKernelGGL(k_seg_gather, dim3(gN), dim3(kSegBlk), 0, st, sg.idx, N, h.dev.atom_x, xa, 0);
  if (v0) HIPCHK(hipMemcpyAsync(vv, v0, sizeof(double) * N3, hipMemcpyHostToDevice, st));
  else hipLaunchKernelGGL(k_seg_gather, dim3(gN), dim3(kSegBlk), 0, st, sg.idx, N, h.dev.atom_v, vv, 0);
  const double dt = d.dt, cc = 0.5 * d.dt * kKickUnit;
  auto tgt = [&](int step) -> const double * { return tgt_of[step] >= 0 ? tg + (size_t)tgt_of[step] * N3 : nullptr; };
  auto sqd = [&](const double *x, int step) {
    if (tgt(step)) hipLaunchKernelGGL(k_seg_sqdist, dim3(1), dim3(kSegBlk), 0, st, N3, x, tgt(step), lp + step);
  };
  // forward: checkpoints at the start of every interval
  double *x = xa, *xn = xb;
  if ((rc = sg.forces(x, F))) return rc;
  sqd(x, 0);
  for (int i = 0; i < nint; i++) {
    const int k0 = i * every, n = std::min(every, nsteps - k0);
    HIPCHK(hipMemcpyAsync(ck + (size_t)2 * i * N3, x, sizeof(double) * N3, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(ck + (size_t)(2 * i + 1) * N3, vv, sizeof(double) * N3, hipMemcpyDeviceToDevice, st));
    for (int j = 1; j <= n; j++) {
      if ((rc = sg.step(x, vv, F, vh, xn, vv, cc, dt))) return rc;
      std::swap(x, xn);
      sqd(x, k0 + j);
    }
  }
  if (x_end) HIPCHK(hipMemcpyAsync(x_end, x, sizeof(double) * N3, hipMemcpyDeviceToHost, st));
  if (v_end) HIPCHK(hipMemcpyAsync(v_end, vv, sizeof(double) * N3, hipMemcpyDeviceToHost, st));
  // reverse, one checkpoint interval at a time
  if (grad) {
    const size_t n3 = N3;
    hipLaunchKernelGGL(k_seg_fill, dim3(seg_grid(n3)), dim3(kSegBlk), 0, st, n3, av, 0.0);
    hipLaunchKernelGGL(k_seg_fill, dim3(seg_grid(3 * (size_t)ni)), dim3(kSegBlk), 0, st, 3 * (size_t)ni, gsum, 0.0);
    if (tgt(nsteps)) hipLaunchKernelGGL(k_seg_dloss, dim3(gN), dim3(kSegBlk), 0, st, N3, x, tgt(nsteps), (double)N, ax);
    else hipLaunchKernelGGL(k_seg_fill, dim3(seg_grid(n3)), dim3(kSegBlk), 0, st, n3, ax, 0.0);
    for (int i = nint - 1; i >= 0; i--) {
      const int k0 = i * every, n = std::min(every, nsteps - k0);
      // positions x_{k0} .. x_{k0+n-1} into frame slots n-1 .. 0 (the
      // reverse order in which the steps are taken back)
      auto slot = [&](int j) { return frames + (size_t)(n - 1 - j) * N3; };
      HIPCHK(hipMemcpyAsync(slot(0), ck + (size_t)2 * i * N3, sizeof(double) * N3, hipMemcpyDeviceToDevice, st));
      if (n > 1) {
        HIPCHK(hipMemcpyAsync(vv, ck + (size_t)(2 * i + 1) * N3, sizeof(double) * N3, hipMemcpyDeviceToDevice, st));
        if ((rc = sg.forces(slot(0), F))) return rc;
        for (int j = 1; j < n; j++)
          if ((rc = sg.step(slot(j - 1), vv, F, vh, slot(j), vv, cc, dt))) return rc;
      }
      for (int j = n - 1; j >= 0; j--) {
        const int stp = k0 + j;
        const double kc = (stp > 0 ? 2.0 : 1.0) * cc;
        double *usl = us + (size_t)(n - 1 - j) * N3;
        hipLaunchKernelGGL(k_seg_back_u, dim3(gN), dim3(kSegBlk), 0, st, N, av, ax, sg.mass, dt, kc, u, usl);
        if ((rc = sg.hvp(slot(j), u, hu))) return rc;
        hipLaunchKernelGGL(k_seg_back_x, dim3(gN), dim3(kSegBlk), 0, st, N3, ax, hu, kc, slot(j), tgt(stp),
                           (double)N);
      }
      if (sg.vi) {
        if ((rc = vi_adjoint_dev(sg.batch(frames, n), nullptr, us, sg.epart, gint, gint + ni, gint + 2 * ni, st)))
          return rc;
        hipLaunchKernelGGL(k_seg_add, dim3(seg_grid(3 * (size_t)ni)), dim3(kSegBlk), 0, st, 3 * ni, gsum, gint);
      }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(g_sigma, gsum, sizeof(double) * ni, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(g_eps, gsum + ni, sizeof(double) * ni, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(g_q, gsum + 2 * ni, sizeof(double) * ni, hipMemcpyDeviceToHost, st));
  }
  std::vector<double> parts((size_t)nsteps + 1, 0.0);
  HIPCHK(hipMemcpyAsync(parts.data(), lp, sizeof(double) * (nsteps + 1), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(h.dev.atom_x, save_x, sizeof(double) * 3 * d.n_atoms, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(h.dev.atom_f, save_f, sizeof(double) * 3 * d.n_atoms, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  // the loss in step order (step 0 last, as its term was added after the
  // segment's)
  double L = 0;
  for (int s = 1; s <= nsteps; s++)
    if (tgt_of[s] >= 0) L += parts[s] / N;
  if (tgt_of[0] >= 0) L += parts[0] / N;
  *loss = L;
  return MDN_OK;
}

}  // extern "C"
