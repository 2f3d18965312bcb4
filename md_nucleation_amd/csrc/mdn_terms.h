// mdn_terms.h -- the per-step operations of one replica, written once over a
// store policy (GStore: global memory; SStore: LDS), plus the workgroup-level
// pair-list rebuild.  Every loop is "for (p = t; p < n; p += nt)" over a flat
// index space, so the same code runs grid-wide (per-call path: t = global
// thread id) or inside one workgroup (batched path: t = threadIdx.x).
// Forces are accumulated with FP64 atomics (global_atomic_add_f64 /
// ds_add_f64 on gfx950); energies per thread in e[U_N].
#pragma once
#include "mdn_device.h"

namespace mdn {

// Local view of the in-domain set: atom order [ion | vapor | gas].
struct LView {
  int n_ion, nv, ng, na_v, na_g;
  int vbase, gbase, n_in;
  int n_gi, n_gv, n_gg;
  __host__ __device__ void set(int nion, int nvin, int ngin, int nav, int nag) {
    n_ion = nion; nv = nvin; ng = ngin; na_v = nav; na_g = nag;
    vbase = nion; gbase = nion + nvin * nav; n_in = gbase + ngin * nag;
  }
};

__device__ inline LView view_from_dyn(const DevReplica &R) {
  LView L;
  L.set(R.sp[0].natoms, R.dyn->n_in_vap, R.dyn->n_in_gas, R.sp[2].natoms, R.sp[1].natoms);
  L.n_gi = R.dyn->n_gi; L.n_gv = R.dyn->n_gv; L.n_gg = R.dyn->n_gg;
  return L;
}

// global-memory store: local index -> (global atom, species, index in molecule)
struct GStore {
  const DevReplica *R;
  LView L;
  typedef double P;  // position and force type of the term routines
  struct H { int g, sp, idx; };
  __device__ H h(int a) const {
    H r;
    if (a < L.n_ion) { r.g = a; r.sp = 0; r.idx = a; return r; }
    if (a < L.gbase) {
      int k = (a - L.vbase) / L.na_v, i = a - L.vbase - k * L.na_v;
      r.g = R->off_vap + (R->in_vap[k] - 1 - R->n_gas) * L.na_v + i; r.sp = 2; r.idx = i; return r;
    }
    int k = (a - L.gbase) / L.na_g, i = a - L.gbase - k * L.na_g;
    r.g = R->off_gas + (R->in_gas[k] - 1) * L.na_g + i; r.sp = 1; r.idx = i; return r;
  }
  __device__ void pos(const H &q, double p[3]) const {
    const double *x = R->atom_x + 3 * (size_t)q.g; p[0] = x[0]; p[1] = x[1]; p[2] = x[2];
  }
  // When a whole wave adds to one atom (a one-atom ion in vi, a gas atom
  // shared by a run of list entries) the wave sums first and one lane adds:
  // 64 atomics on one address would serialise in the atomic unit.
  __device__ void fadd(const H &q, double fx, double fy, double fz) const {
    const int g0 = __builtin_amdgcn_readfirstlane(q.g);
    if (__ballot(1) == ~0ull && __ballot(q.g == g0) == ~0ull) {
      fx = wave_sum(fx);
      fy = wave_sum(fy);
      fz = wave_sum(fz);
      if ((threadIdx.x & 63) != 0) return;
    }
    double *f = R->atom_f + 3 * (size_t)q.g;
    atomicAdd(f, fx); atomicAdd(f + 1, fy); atomicAdd(f + 2, fz);
  }
  __device__ double w(const H &q) const { return R->atom_w[q.g]; }
  __device__ double q(const H &h_) const { return R->sp[h_.sp].charge[h_.idx]; }
  __device__ int type(const H &h_) const { return R->sp[h_.sp].type[h_.idx]; }
  // pair_coeff table of setCrossPotentials (setCrossPotential.cpp:4-23)
  template <typename R2>
  __device__ void coeff(const DevReplica &Rr, const H &a, const H &b, R2 &c0, R2 &c1) const {
    const double *p = Rr.pair_coeff + 2 * ((size_t)type(a) * Rr.n_types + type(b));
    c0 = (R2)p[0];
    c1 = (R2)p[1];
  }
};

// forward-mode store over global memory: the positions are x + eps u as
// dual numbers, and each force's derivative along u (d F / d eps, the
// Hessian of -U applied to u) is added to hu; the forces themselves are not
// written.  Pair lists and blend weights are held as they stand, as in one
// force evaluation.
struct JStore : GStore {
  typedef Dual P;
  const double *u;  // [n_atoms][3] direction
  double *hu;       // [n_atoms][3] derivative of the forces along u
  __device__ void pos(const H &q, Dual p[3]) const {
    const double *x = R->atom_x + 3 * (size_t)q.g, *d = u + 3 * (size_t)q.g;
    p[0] = Dual(x[0], d[0]); p[1] = Dual(x[1], d[1]); p[2] = Dual(x[2], d[2]);
  }
  __device__ void fadd(const H &q, const Dual &fx, const Dual &fy, const Dual &fz) const {
    double *f = hu + 3 * (size_t)q.g;
    atomicAdd(f, fx.d); atomicAdd(f + 1, fy.d); atomicAdd(f + 2, fz.d);
  }
};

// JStore that keeps the derivative of the force on one atom (global index
// `row`) in registers and drops the rest: the i-centric form of a pair
// term's derivative, each pair evaluated once from each side
struct JRowStore : JStore {
  int row;
  double *acc;  // [3]
  __device__ JRowStore(const JStore &b, int r, double *a) : JStore(b), row(r), acc(a) {}
  __device__ void fadd(const H &q, const Dual &fx, const Dual &fy, const Dual &fz) const {
    if (q.g == row) { acc[0] += fx.d; acc[1] += fy.d; acc[2] += fz.d; }
  }
};

// the same for the forces themselves
struct GRowStore : GStore {
  int row;
  double *acc;  // [3]
  __device__ GRowStore(const GStore &b, int r, double *a) : GStore(b), row(r), acc(a) {}
  __device__ void fadd(const H &q, double fx, double fy, double fz) const {
    if (q.g == row) { acc[0] += fx; acc[1] += fy; acc[2] += fz; }
  }
};
__device__ inline GRowStore row_store(const GStore &s, int r, double *a) { return GRowStore(s, r, a); }
__device__ inline JRowStore row_store(const JStore &s, int r, double *a) { return JRowStore(s, r, a); }
// one atom's summed row added to the store's output (one lane)
__device__ inline void row_flush(const GStore &s, const GStore::H &h, const double *a) {
  double *f = s.R->atom_f + 3 * (size_t)h.g;
  atomicAdd(f, a[0]); atomicAdd(f + 1, a[1]); atomicAdd(f + 2, a[2]);
}
__device__ inline void row_flush(const JStore &s, const GStore::H &h, const double *a) {
  double *f = s.hu + 3 * (size_t)h.g;
  atomicAdd(f, a[0]); atomicAdd(f + 1, a[1]); atomicAdd(f + 2, a[2]);
}

__device__ inline void pc(const DevReplica &R, int t1, int t2, double &c0, double &c1) {
  const double *p = R.pair_coeff + 2 * ((size_t)t1 * R.n_types + t2);
  c0 = p[0]; c1 = p[1];
}

// one LJ / LJcoul atom pair: F_a += fp*d, F_b -= fp*d (d = x_a - x_b)
template <typename Rl, bool COUL, class S>
__device__ inline void pair_ab(const DevReplica &R, const S &s, const typename S::H &ha, const typename S::H &hb,
                               bool weighted, bool eflag, double &e) {
  typedef typename S::P P;
  P xa[3], xb[3];
  s.pos(ha, xa);
  s.pos(hb, xb);
  P dx = xa[0] - xb[0], dy = xa[1] - xb[1], dz = xa[2] - xb[2];
  Rl c0, c1;
  s.template coeff<Rl>(R, ha, hb, c0, c1);
  Rl rx = (Rl)dx, ry = (Rl)dy, rz = (Rl)dz;
  Rl rsq = rx * rx + ry * ry + rz * rz;
  Rl w = weighted ? (Rl)(s.w(ha) * s.w(hb)) : Rl(1);
  Rl fp;
  if (COUL)
    fp = ljcoul_fpair<Rl>(rsq, c0, c1, (Rl)(R.qqrd2e * s.q(ha) * s.q(hb)), w, eflag, e);
  else
    fp = lj_fpair<Rl>(rsq, c0, c1, w, eflag, e);
  P fx = (P)(fp * rx), fy = (P)(fp * ry), fz = (P)(fp * rz);
  s.fadd(ha, fx, fy, fz);
  s.fadd(hb, -fx, -fy, -fz);
}

// ---- gi: PotentialGasIon::compute, GasIon.cpp:4-38 (list entries (gas
// molecule, ion atom); every gas atom of the molecule; unweighted) ----
template <typename Rl, class S>
__device__ void term_gi(const DevReplica &R, const S &s, int t, int nt, bool eflag, double *e) {
  const LView &L = s.L;
  int n = L.n_gi * L.na_g;
  for (int p = t; p < n; p += nt) {
    int ent = p / L.na_g, ii = p - ent * L.na_g;
    int k = R.gi[2 * ent], j = R.gi[2 * ent + 1];
    pair_ab<Rl, false>(R, s, s.h(L.gbase + k * L.na_g + ii), s.h(j), false, eflag, e[U_GI]);
  }
}

// ---- vi: PotentialVaporIon::compute, VaporIon.cpp:5-43 (all in-domain
// vapor atoms x all ion atoms, weighted) ----
template <typename Rl, class S>
__device__ void term_vi(const DevReplica &R, const S &s, int t, int nt, bool eflag, double *e) {
  const LView &L = s.L;
  int n = L.nv * L.na_v * L.n_ion;
  for (int p = t; p < n; p += nt) {
    int a = p / L.n_ion, j = p - a * L.n_ion;
    pair_ab<Rl, true>(R, s, s.h(L.vbase + a), s.h(j), true, eflag, e[U_VI]);
  }
}

// ---- gg: gas-gas LJ over the COM list, weighted w_a w_b (the evident
// intent of GasGas.cpp:4-63, whose list indexes loop counters) ----
template <typename Rl, class S>
__device__ void term_gg(const DevReplica &R, const S &s, int t, int nt, bool eflag, double *e) {
  const LView &L = s.L;
  int nn = L.na_g * L.na_g, n = L.n_gg * nn;
  for (int p = t; p < n; p += nt) {
    int ent = p / nn, r = p - ent * nn, ii = r / L.na_g, jj = r - ii * L.na_g;
    int k = R.gg[2 * ent], l = R.gg[2 * ent + 1];
    pair_ab<Rl, false>(R, s, s.h(L.gbase + k * L.na_g + ii), s.h(L.gbase + l * L.na_g + jj), true, eflag,
                       e[U_GG]);
  }
}

// ---- gv: PotentialVaporGas::compute, VaporGas.cpp:5-44 (unweighted; the
// inner bound is the gas atom count, VaporGas.cpp:15) ----
template <typename Rl, class S>
__device__ void term_gv(const DevReplica &R, const S &s, int t, int nt, bool eflag, double *e) {
  const LView &L = s.L;
  int nj = L.na_g < L.na_v ? L.na_g : L.na_v;
  int nn = L.na_g * nj, n = L.n_gv * nn;
  for (int p = t; p < n; p += nt) {
    int ent = p / nn, r = p - ent * nn, ii = r / nj, jj = r - ii * nj;
    int k = R.gv[2 * ent], l = R.gv[2 * ent + 1];
    pair_ab<Rl, false>(R, s, s.h(L.gbase + k * L.na_g + ii), s.h(L.vbase + l * L.na_v + jj), false, eflag,
                       e[U_VG]);
  }
}

// ---- vv: PotentialVaporVapor::compute, VaporVapor.cpp:5-45 (all atom pairs
// of distinct in-domain vapor molecules, weighted) ----
template <typename Rl, class S>
__device__ void term_vv(const DevReplica &R, const S &s, int t, int nt, bool eflag, double *e) {
  const LView &L = s.L;
  int nva = L.nv * L.na_v;
  // pairs (a, b) with molecule(a) < molecule(b): iterate the strictly upper
  // block triangle row by row
  long long n = (long long)nva * nva;
  for (long long p = t; p < n; p += nt) {
    int a = (int)(p / nva), b = (int)(p - (long long)a * nva);
    if (a / L.na_v >= b / L.na_v) continue;
    pair_ab<Rl, true>(R, s, s.h(L.vbase + a), s.h(L.vbase + b), true, eflag, e[U_VV]);
  }
}

// the derivative of vv along u, i-centric: one wave per in-domain vapor atom
// a (waves w of nw), its lanes over every vapor atom of another molecule;
// the wave's sum reaches hu with one atomic per component (no atomic per
// pair)
__device__ inline void term_vv_jvp_rows(const DevReplica &R, const JStore &s, int w, int nw, int lane) {
  const LView &L = s.L;
  const int nva = L.nv * L.na_v;
  for (int a = w; a < nva; a += nw) {
    double acc[3] = {0, 0, 0}, e = 0;
    const JStore::H ha = s.h(L.vbase + a);
    const JRowStore rs(s, ha.g, acc);
    for (int b = lane; b < nva; b += 64) {
      if (a / L.na_v == b / L.na_v) continue;
      pair_ab<Dual, true>(R, rs, ha, rs.h(L.vbase + b), true, false, e);
    }
    for (int q = 0; q < 3; q++) acc[q] = wave_sum_full(acc[q]);
    if (lane == 0) {
      double *h = s.hu + 3 * (size_t)ha.g;
      atomicAdd(h, acc[0]); atomicAdd(h + 1, acc[1]); atomicAdd(h + 2, acc[2]);
    }
  }
}

// ---- Efield: PotentialEfield::compute, electricField.cpp:54-64 ----
template <class S>
__device__ void term_efield(const DevReplica &R, const S &s, int t, int nt) {
  for (int i = t; i < s.L.n_ion; i += nt) {
    typename S::H h = s.h(i);
    double c = 5.55e-5 * s.q(h);
    s.fadd(h, c * R.efield[0], c * R.efield[1], c * R.efield[2]);
  }
}

// bonded terms of one species' molecule whose first local atom is `base`
template <class S>
__device__ inline void bond_term(const DevReplica &R, const S &s, int base, const int *b, bool eflag, double &e) {
  typedef typename S::P P;
  typename S::H hi = s.h(base + b[0]), hj = s.h(base + b[1]);
  P xi[3], xj[3];
  s.pos(hi, xi); s.pos(hj, xj);
  P d[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
  P fb = bond_fb<P>(d[0] * d[0] + d[1] * d[1] + d[2] * d[2], R.btype[2 * b[2]], R.btype[2 * b[2] + 1], eflag, e);
  s.fadd(hi, fb * d[0], fb * d[1], fb * d[2]);
  s.fadd(hj, -fb * d[0], -fb * d[1], -fb * d[2]);
}
template <class S>
__device__ inline void angle_term(const DevReplica &R, const S &s, int base, const int *a, bool eflag, double &e) {
  typedef typename S::P P;
  typename S::H hi = s.h(base + a[0]), hj = s.h(base + a[1]), hk = s.h(base + a[2]);
  P xi[3], xj[3], xk[3], f1[3], f3[3];
  s.pos(hi, xi); s.pos(hj, xj); s.pos(hk, xk);
  P d1[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
  P d2[3] = {xk[0] - xj[0], xk[1] - xj[1], xk[2] - xj[2]};
  angle_f<P>(d1, d2, R.ctype[2 * a[3]], R.ctype[2 * a[3] + 1], eflag, e, f1, f3);
  s.fadd(hi, f1[0], f1[1], f1[2]);
  s.fadd(hj, -(f1[0] + f3[0]), -(f1[1] + f3[1]), -(f1[2] + f3[2]));
  s.fadd(hk, f3[0], f3[1], f3[2]);
}
template <class S>
__device__ inline void dihedral_term(const DevReplica &R, const S &s, int base, const int *d, bool eflag,
                                     double &e) {
  typedef typename S::P P;
  typename S::H h1 = s.h(base + d[0]), h2 = s.h(base + d[1]), h3 = s.h(base + d[2]), h4 = s.h(base + d[3]);
  P x1[3], x2[3], x3[3], x4[3], f1[3], f2[3], f3[3], f4[3];
  s.pos(h1, x1); s.pos(h2, x2); s.pos(h3, x3); s.pos(h4, x4);
  P vb1[3] = {x1[0] - x2[0], x1[1] - x2[1], x1[2] - x2[2]};
  P vb2[3] = {x3[0] - x2[0], x3[1] - x2[1], x3[2] - x2[2]};
  P vb3[3] = {x4[0] - x3[0], x4[1] - x3[1], x4[2] - x3[2]};
  dihedral_f<P>(vb1, vb2, vb3, R.dcoeff + 15 * d[4], R.dmulti[d[4]], eflag, e, f1, f2, f3, f4);
  s.fadd(h1, f1[0], f1[1], f1[2]);
  s.fadd(h2, f2[0], f2[1], f2[2]);
  s.fadd(h3, f3[0], f3[1], f3[2]);
  s.fadd(h4, f4[0], f4[1], f4[2]);
}

// ---- AMBER ion: PotentialAMBER::compute, compute_AMBER.cpp:10-253
// (non-excluded pairs from initialAMBER, unweighted LJcoul; bonds, angles,
// dihedrals; all into Uion) ----
template <class S>
__device__ void term_amber_bonded(const DevReplica &R, const S &s, int t, int nt, bool eflag, double *e) {
  const DevSpecies &I = R.sp[0];
  // one index space over bonds, angles and dihedrals, so that each thread
  // takes about one term
  const int nb = I.nbond, nba = nb + I.nangle, n = nba + I.ndih;
  for (int p = t; p < n; p += nt) {
    if (p < nb) bond_term(R, s, 0, I.bond + 3 * p, eflag, e[U_ION]);
    else if (p < nba) angle_term(R, s, 0, I.angle + 4 * (p - nb), eflag, e[U_ION]);
    else dihedral_term(R, s, 0, I.dih + 5 * (p - nba), eflag, e[U_ION]);
  }
}
// The AMBER non-bonded pairs (compute_AMBER.cpp:14-50 over the list of
// initialAMBER) over global memory, i-centric: wave w of nw takes ion atoms
// i = w, w + nw, ..., its lanes the partners j that row i of the exclusion
// bitmask allows; the force on i is summed in registers and added with one
// atomic per component (a pair list walked pair by pair would add six
// atomics per pair onto the ion's few hundred force slots).  Every pair is
// met from both sides, so its energy is halved.
template <typename Rl, class S>
__device__ void term_amber_long_rows(const DevReplica &R, const S &s, int w, int nw, int lane, bool eflag,
                                     double *e) {
  const int n = R.sp[0].natoms, W = R.amber_words;
  double eh = 0;
  for (int i = w; i < n; i += nw) {
    double acc[3] = {0, 0, 0};
    const typename S::H hi = s.h(i);
    const auto rs = row_store(s, hi.g, acc);
    const uint32_t *row = R.amber_allow + (size_t)i * W;
    for (int j = lane; j - lane < n; j += 64)
      if (j < n && ((row[j >> 5] >> (j & 31)) & 1u)) pair_ab<Rl, true>(R, rs, hi, rs.h(j), false, eflag, eh);
    for (int q = 0; q < 3; q++) acc[q] = wave_sum_full(acc[q]);
    if (lane == 0) row_flush(s, hi, acc);
  }
  if (eflag) e[U_ION] += 0.5 * eh;
}

// ---- vapor intra: PotentialVaporIntra::compute, compute_intra.cpp:45-245
// (bonds + angles -> Ugas, dihedrals -> Uion, as labelled there) ----
template <class S>
__device__ void term_vintra(const DevReplica &R, const S &s, int t, int nt, bool eflag, double *e) {
  const LView &L = s.L;
  const DevSpecies &V = R.sp[2];
  int per = V.nbond + V.nangle + V.ndih;
  int n = L.nv * per;
  for (int p = t; p < n; p += nt) {
    int k = p / per, r = p - k * per;
    int base = L.vbase + k * L.na_v;
    if (r < V.nbond) bond_term(R, s, base, V.bond + 3 * r, eflag, e[U_GAS]);
    else if (r < V.nbond + V.nangle) angle_term(R, s, base, V.angle + 4 * (r - V.nbond), eflag, e[U_GAS]);
    else dihedral_term(R, s, base, V.dih + 5 * (r - V.nbond - V.nangle), eflag, e[U_ION]);
  }
}

// ---- gas intra: PotentialGasIntra::compute, compute_intra.cpp:14-41 ----
template <class S>
__device__ void term_gintra(const DevReplica &R, const S &s, int t, int nt, bool eflag, double *e) {
  const LView &L = s.L;
  const DevSpecies &G = R.sp[1];
  int n = L.ng * G.nbond;
  for (int p = t; p < n; p += nt) {
    int k = p / G.nbond, r = p - k * G.nbond;
    bond_term(R, s, L.gbase + k * L.na_g, G.bond + 3 * r, eflag, e[U_GAS]);
  }
}

// ---- Born-Mayer-Huggins NaCl: PotentialBorn::compute, compute_born.cpp:10-56
// with the coefficient table of potentialBMHFT.hpp:13-41 indexed by ion-local
// type (0 = Na, 1 = Cl) ----
__device__ inline void born_coeff(int t1, int t2, double B[5]) {
  int k = t1 + t2;  // 0 NaNa, 1 NaCl, 2 ClCl
  double A = k == 0 ? 25.4435 : (k == 1 ? 20.3548 : 15.2661);
  double C = k == 0 ? 101.1719 : (k == 1 ? 674.4793 : 6985.6786);
  double D = k == 0 ? 48.1771 : (k == 1 ? 837.077 : 14031.5785);
  B[0] = A * 3.1546 / 4.184;
  B[1] = 6 * C / 4.184;
  B[2] = 8 * D / 4.184;
  B[3] = k == 0 ? 2.340 : (k == 1 ? 2.755 : 3.170);
  B[4] = 3.1546;
}
template <class S>
__device__ void term_born(const DevReplica &R, const S &s, int t, int nt, bool eflag, double *e) {
  int is = s.L.n_ion, n = is * is;
  for (int p = t; p < n; p += nt) {
    int i = p / is, j = p - i * is;
    if (i >= j) continue;
    typename S::H hi = s.h(i), hj = s.h(j);
    double xi[3], xj[3];
    s.pos(hi, xi); s.pos(hj, xj);
    double dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
    double rsq = dx * dx + dy * dy + dz * dz;
    double r2inv = 1 / rsq, r6inv = r2inv * r2inv * r2inv, r = sqrt(rsq);
    int t1 = s.type(hi) - R.ion_type_base, t2 = s.type(hj) - R.ion_type_base;
    t1 = t1 < 0 ? 0 : (t1 > 1 ? 1 : t1);
    t2 = t2 < 0 ? 0 : (t2 > 1 ? 1 : t2);
    double B[5];
    born_coeff(t1, t2, B);
    double rexp = exp(B[4] * (B[3] - r));
    double f1 = B[4] * B[0] * r * rexp, f2 = -B[1] * r6inv, f3 = -B[2] * r6inv * r2inv;
    double fc = R.qqrd2e * s.q(hi) * s.q(hj) * sqrt(r2inv);
    double fp = (f1 + f2 + f3 + fc) * r2inv;
    s.fadd(hi, fp * dx, fp * dy, fp * dz);
    s.fadd(hj, -fp * dx, -fp * dy, -fp * dz);
    if (eflag) e[U_ION] += fc + rexp * B[0] - B[1] / 6.0 * r6inv - B[2] / 8.0 * r6inv * r2inv;
  }
}

// ---- Tersoff (Tersoff/Tersoff.cpp:13-297, Si constants of
// potentialTersoff.hpp:15-39) with its neighbour list built; forces and
// energy converted eV -> kcal/mol by 23.06 on its own contribution ----
struct TersP {
  double gamma = 1.0, lamda3 = 0, c2 = 100390.0 * 100390.0, d2 = 16.218 * 16.218, h0 = -0.59826, n = 0.78734,
         beta = 1.0999e-6, lamda2 = 1.7322, B = 471.18, R = 2.85, D = 0.15, lamda1 = 2.4799, A = 1830.8;
};
__device__ inline double t_fC(const TersP &p, double r) {
  if (r < p.R - p.D) return 1;
  if (r > p.R + p.D) return 0;
  return 0.5 * (1 - sin(kPI * 0.5 * (r - p.R) / p.D));
}
__device__ inline double t_fC_d(const TersP &p, double r) {
  if (r < p.R - p.D || r > p.R + p.D) return 0;
  return -0.25 * kPI / p.D * cos(kPI * 0.5 * (r - p.R) / p.D);
}
__device__ inline double t_g(const TersP &p, double ct) {
  double h = p.h0 - ct;
  return p.gamma * (1.0 + p.c2 / p.d2 - p.c2 / (p.d2 + h * h));
}
__device__ inline double t_g_d(const TersP &p, double ct) {
  double h = p.h0 - ct, den = 1 / (p.d2 + h * h);
  return p.gamma * (-2.0 * p.c2 * h) * den * den;
}
__device__ inline double t_bij(const TersP &p, double z) {
  if (z == 0) return 1.0;
  return pow(1.0 + pow(p.beta * z, p.n), -1.0 / (2.0 * p.n));
}
__device__ inline double t_bij_d(const TersP &p, double z) {
  double tn = pow(p.beta * z, p.n);
  return -0.5 * pow(1.0 + tn, -1.0 - 1.0 / (2.0 * p.n)) * tn / z;
}
template <class S>
__device__ void term_tersoff(const DevReplica &R, const S &s, int t, int nt, bool eflag, double *e) {
  TersP P;
  const double cut = P.R + P.D, cut2 = cut * cut, sc = 23.06;
  int is = s.L.n_ion;
  for (int i = t; i < is; i += nt) {
    typename S::H hi = s.h(i);
    double xi[3];
    s.pos(hi, xi);
    double U = 0;
    for (int j = 0; j < is; j++) {
      if (j == i) continue;
      typename S::H hj = s.h(j);
      double xj[3];
      s.pos(hj, xj);
      double d[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
      double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (rsq > cut2) continue;
      double r = sqrt(rsq);
      double fR = P.A * exp(-P.lamda1 * r), fR_d = -P.A * P.lamda1 * exp(-P.lamda1 * r);
      double fC = t_fC(P, r), fC_d = t_fC_d(P, r);
      double fA = -P.B * exp(-P.lamda2 * r), fA_d = P.B * P.lamda2 * exp(-P.lamda2 * r);
      // repulsive pair
      double fp = -(fC * fR_d + fR * fC_d) / r * 0.5;
      U += fC * fR * 0.5;
      // bond order
      double zeta = 0;
      for (int k = 0; k < is; k++) {
        if (k == i || k == j) continue;
        double xk[3];
        s.pos(s.h(k), xk);
        double d2[3] = {xi[0] - xk[0], xi[1] - xk[1], xi[2] - xk[2]};
        double rsq2 = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
        if (rsq2 > cut2) continue;
        double r2 = sqrt(rsq2);
        double ct = (d[0] * d2[0] + d[1] * d2[1] + d[2] * d2[2]) / (r * r2);
        zeta += exp(P.lamda3 * P.lamda3 * P.lamda3 * (r - r2) * (r - r2) * (r - r2)) * t_g(P, ct) * t_fC(P, r2);
      }
      double bij = t_bij(P, zeta);
      double fz = bij * (fA * fC_d + fA_d * fC) / r;
      fp += -0.5 * fz;
      U += bij * fC * fA * 0.5;
      s.fadd(hi, sc * fp * d[0], sc * fp * d[1], sc * fp * d[2]);
      s.fadd(hj, -sc * fp * d[0], -sc * fp * d[1], -sc * fp * d[2]);
      if (zeta == 0) continue;
      double pre = fA * fC * t_bij_d(P, zeta);
      double rij_hat[3] = {d[0] / r, d[1] / r, d[2] / r};
      double fi[3] = {0, 0, 0}, fj[3] = {0, 0, 0};
      for (int k = 0; k < is; k++) {
        if (k == i || k == j) continue;
        typename S::H hk = s.h(k);
        double xk[3];
        s.pos(hk, xk);
        double d2[3] = {xi[0] - xk[0], xi[1] - xk[1], xi[2] - xk[2]};
        double rsq2 = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
        if (rsq2 > cut2) continue;
        double r2 = sqrt(rsq2);
        double rik_hat[3] = {d2[0] / r2, d2[1] / r2, d2[2] / r2};
        double fc = t_fC(P, r2), dfc = t_fC_d(P, r2);
        double ct = rij_hat[0] * rik_hat[0] + rij_hat[1] * rik_hat[1] + rij_hat[2] * rik_hat[2];
        double g = t_g(P, ct), g_d = t_g_d(P, ct);
        double fk[3];
        for (int q = 0; q < 3; q++) {
          double dcdrj = (rik_hat[q] - ct * rij_hat[q]) / r;
          double dcdrk = (rij_hat[q] - ct * rik_hat[q]) / r2;
          double dcdri = -dcdrj - dcdrk;
          fi[q] += 0.5 * pre * (-dfc * g * rik_hat[q] + fc * g_d * dcdri);
          fj[q] += 0.5 * pre * (fc * g_d * dcdrj);
          fk[q] = 0.5 * pre * (dfc * g * rik_hat[q] + fc * g_d * dcdrk);
        }
        s.fadd(hk, sc * fk[0], sc * fk[1], sc * fk[2]);
      }
      s.fadd(hi, sc * fi[0], sc * fi[1], sc * fi[2]);
      s.fadd(hj, sc * fj[0], sc * fj[1], sc * fj[2]);
    }
    if (eflag) e[U_ION] += U * sc;
  }
}

// ---- Stillinger-Weber (StillingerWeber/StillingerWeber.cpp:11-198, Si
// constants of potentialSW.hpp:13-37); same list / unit treatment ----
struct SWP {
  double c1, c2, c3, c4, c5, c6, sigma, sigma_gamma, lambda_epsilon, lambda_epsilon2, costheta0, p, q, cut;
  __device__ SWP() {
    double epsilon = 2.1683, sg = 2.0951, a = 1.80, lambda = 21.0, gamma = 1.20;
    double A = 7.049556277, B = 0.6022245584;
    p = 4.0; q = 0.0; costheta0 = -0.333333333333; sigma = sg;
    c1 = A * epsilon * p * B * pow(sg, p);
    c2 = A * epsilon * q * pow(sg, q);
    c3 = A * epsilon * B * pow(sg, p + 1);
    c4 = A * epsilon * pow(sg, q + 1);
    c5 = A * epsilon * B * pow(sg, p);
    c6 = A * epsilon * pow(sg, q);
    sigma_gamma = sg * gamma; lambda_epsilon = lambda * epsilon; lambda_epsilon2 = 2 * lambda * epsilon;
    cut = a * sg;
  }
};
template <class S>
__device__ void term_sw(const DevReplica &R, const S &s, int t, int nt, bool eflag, double *e) {
  SWP P;
  const double cut2 = P.cut * P.cut, sc = 23.06;
  int is = s.L.n_ion;
  for (int i = t; i < is; i += nt) {
    typename S::H hi = s.h(i);
    double xi[3];
    s.pos(hi, xi);
    double U = 0, fi[3] = {0, 0, 0};
    for (int j = 0; j < is; j++) {
      if (j == i) continue;
      typename S::H hj = s.h(j);
      double xj[3];
      s.pos(hj, xj);
      double d[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
      double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (rsq > cut2) continue;
      // two-body, StillingerWeber.cpp:104-116
      double r = sqrt(rsq), rinvsq = 1.0 / rsq;
      double rp = pow(r, -P.p), rq = pow(r, -P.q);
      double rainv = 1.0 / (r - P.cut), rainvsq = rainv * rainv * r, expsrainv = exp(P.sigma * rainv);
      U += (P.c5 * rp - P.c6 * rq) * expsrainv;
      double fp = (P.c1 * rp - P.c2 * rq + (P.c3 * rp - P.c4 * rq) * rainvsq) * expsrainv * rinvsq * 0.5;
      for (int q = 0; q < 3; q++) fi[q] += fp * d[q];
      s.fadd(hj, -sc * fp * d[0], -sc * fp * d[1], -sc * fp * d[2]);
      // three-body over k != j within the cutoff, StillingerWeber.cpp:120-168
      double fj[3] = {0, 0, 0};
      for (int k = 0; k < is; k++) {
        if (k == i || k == j) continue;
        typename S::H hk = s.h(k);
        double xk[3];
        s.pos(hk, xk);
        double d2[3] = {xi[0] - xk[0], xi[1] - xk[1], xi[2] - xk[2]};
        double rsq2 = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
        if (rsq2 > cut2) continue;
        double r1 = r, rinvsq1 = rinvsq, rainv1 = 1.0 / (r1 - P.cut);
        double gs1 = P.sigma_gamma * rainv1, gsq1 = gs1 * rainv1 / r1, eg1 = exp(gs1);
        double r2 = sqrt(rsq2), rinvsq2 = 1.0 / rsq2, rainv2 = 1.0 / (r2 - P.cut);
        double gs2 = P.sigma_gamma * rainv2, gsq2 = gs2 * rainv2 / r2, eg2 = exp(gs2);
        double rinv12 = 1.0 / (r1 * r2);
        double cs = (d[0] * d2[0] + d[1] * d2[1] + d[2] * d2[2]) * rinv12;
        double delcs = cs - P.costheta0, delcssq = delcs * delcs;
        double facexp = eg1 * eg2, facrad = P.lambda_epsilon * facexp * delcssq;
        double frad1 = facrad * gsq1, frad2 = facrad * gsq2;
        double facang = P.lambda_epsilon2 * facexp * delcs, facang12 = rinv12 * facang;
        double csfacang = cs * facang, csfac1 = rinvsq1 * csfacang, csfac2 = rinvsq2 * csfacang;
        double tj[3], tk[3];
        for (int q = 0; q < 3; q++) {
          tj[q] = d[q] * (frad1 + csfac1) - d2[q] * facang12;
          tk[q] = d2[q] * (frad2 + csfac2) - d[q] * facang12;
          fi[q] += (tj[q] + tk[q]) * 0.5;
          fj[q] -= tj[q] * 0.5;
        }
        U += facrad;
        s.fadd(hk, -sc * tk[0] * 0.5, -sc * tk[1] * 0.5, -sc * tk[2] * 0.5);
      }
      s.fadd(hj, sc * fj[0], sc * fj[1], sc * fj[2]);
    }
    s.fadd(hi, sc * fi[0], sc * fi[1], sc * fi[2]);
    if (eflag) e[U_ION] += U * 0.5 * sc;
  }
}

// all registered potentials in the registration order of MD::setPotential
// (setPotential.cpp:13-108): InterInter gi, vi, gg, gv, vv, Efield; then
// IntraInter Tersoff, SW, AMBER, BMH, vapor intra, gas intra.
template <typename Rl, class S>
__device__ void all_terms(const DevReplica &R, const S &s, uint32_t bits, int t, int nt, bool eflag, double *e) {
  // Each term's work items start where the previous term's ended (modulo
  // the grid, in whole waves), so the terms spread over the grid side by
  // side instead of all starting on thread 0: a thread then takes about
  // (total items / threads) items rather than one of every term, and the
  // kernel's latency is that of its slowest items, not of their sum.  A
  // wave still holds items of one term only (GStore::fadd's wave sums rely
  // on it).
  const LView &L = s.L;
  int off = 0;
  auto at = [&](int items) {
    const int tt = t - off;
    off = (int)((off + ((long long)items + 63) / 64 * 64) % nt);
    return tt < 0 ? tt + nt : tt;
  };
  const DevSpecies &I = R.sp[0], &G = R.sp[1], &V = R.sp[2];
  const int nj = L.na_g < L.na_v ? L.na_g : L.na_v;
  if (bits & B_GI) term_gi<Rl>(R, s, at(L.n_gi * L.na_g), nt, eflag, e);
  if (bits & B_VI) term_vi<Rl>(R, s, at(L.nv * L.na_v * L.n_ion), nt, eflag, e);
  if (bits & B_GG) term_gg<Rl>(R, s, at(L.n_gg * L.na_g * L.na_g), nt, eflag, e);
  if (bits & B_GV) term_gv<Rl>(R, s, at(L.n_gv * L.na_g * nj), nt, eflag, e);
  if (bits & B_VV) term_vv<Rl>(R, s, at(0), nt, eflag, e);
  if (bits & B_EFIELD) term_efield(R, s, at(L.n_ion), nt);
  if (bits & B_TERSOFF) term_tersoff(R, s, at(0), nt, eflag, e);
  if (bits & B_SW) term_sw(R, s, at(0), nt, eflag, e);
  if (bits & B_AMBER) {
    const int ta = at(64 * I.natoms);
    term_amber_long_rows<Rl>(R, s, ta >> 6, nt >> 6, ta & 63, eflag, e);
    term_amber_bonded(R, s, at(I.nbond + I.nangle + I.ndih), nt, eflag, e);
  }
  if (bits & B_BMH) term_born(R, s, at(0), nt, eflag, e);
  if (bits & B_VINTRA) term_vintra(R, s, at(L.nv * (V.nbond + V.nangle + V.ndih)), nt, eflag, e);
  if (bits & B_GINTRA) term_gintra(R, s, at(L.ng * G.nbond), nt, eflag, e);
}

// ------------------------------------------------------------------------
// SHAKE of one molecule (MD::shake, shake.cpp:44-92: positions only,
// Gauss-Seidel over the rigid pairs until (r^2 - d^2)^2 <= 1e-20 for all),
// with an iteration cap.  x: current positions (updated), r: positions before
// the drift.  Returns iterations used, negative when the cap was hit.
template <class XA, class RA>
__device__ int shake_molecule_generic(const DevReplica &R, const DevSpecies &Sp, XA x, RA r, int max_iter) {
  const double tolerance = 1e-20, rho_zero = 0.01;
  int it = 0;
  bool converge = false;
  while (!converge) {
    if (it >= max_iter) return -it;
    it++;
    converge = true;
    for (int k = 0; k < Sp.nrigid; k++) {
      int i = Sp.rigid[3 * k], j = Sp.rigid[3 * k + 1], ty = Sp.rigid[3 * k + 2];
      double dk = R.btype[2 * ty + 1];
      double rk[3], rn[3];
      for (int q = 0; q < 3; q++) { rk[q] = r(i, q) - r(j, q); rn[q] = x(i, q) - x(j, q); }
      double rk2 = rk[0] * rk[0] + rk[1] * rk[1] + rk[2] * rk[2];
      double rn2 = rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2];
      double dot = rk[0] * rn[0] + rk[1] * rn[1] + rk[2] * rn[2];
      if (R.shake != 2) {  // mode 1 tests each constraint before its update
        double d0 = rn2 - dk * dk;
        if (d0 * d0 > tolerance) converge = false;
      }
      double mred_inv = Sp.invmass[i] + Sp.invmass[j];
      // mode 2: shake.cpp:67 multiplies by the inverse reduced mass; mode 1
      // divides by it (textbook SHAKE), through the precomputed 0.5 / mred_inv
      double ramk = R.shake == 2 ? (dk * dk - rn2) / 2 * mred_inv : (dk * dk - rn2) * (0.5 / mred_inv);
      if (dot > 1e-10) ramk /= dot;
      else ramk = 0.5 * rho_zero + ramk / (rk2 * rho_zero) * mred_inv;
      double ci = ramk * Sp.invmass[i], cj = ramk * Sp.invmass[j];
      for (int q = 0; q < 3; q++) { x(i, q) += rk[q] * ci; x(j, q) -= rk[q] * cj; }
      for (int q = 0; q < 3; q++) rn[q] = x(i, q) - x(j, q);
      rn2 = rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2];
      double dr = rn2 - dk * dk;
      if (!(dr * dr <= 1e30)) return -it;  // diverged (or non-finite)
      if (R.shake == 2 && dr * dr > tolerance) converge = false;
    }
  }
  return it;
}

// Same iteration with the constraint data (atoms, old bond vector, d^2,
// inverse masses) read once into registers: up to kMaxRigid constraints per
// molecule (methanol under SHAKE H has 4, water 2); more take the generic loop.
constexpr int kMaxRigid = 4;
template <class XA, class RA>
__device__ int shake_molecule(const DevReplica &R, const DevSpecies &Sp, XA x, RA r, int max_iter) {
  const int nr = Sp.nrigid;
  if (nr > kMaxRigid) return shake_molecule_generic(R, Sp, x, r, max_iter);
  const double tolerance = 1e-20, rho_zero = 0.01;
  int ci_[kMaxRigid], cj_[kMaxRigid];
  double rk[kMaxRigid][3], dk2[kMaxRigid], mi[kMaxRigid], mj[kMaxRigid], hm[kMaxRigid];
#pragma unroll
  for (int k = 0; k < kMaxRigid; k++) {
    if (k >= nr) break;
    int i = Sp.rigid[3 * k], j = Sp.rigid[3 * k + 1];
    double dk = R.btype[2 * Sp.rigid[3 * k + 2] + 1];
    ci_[k] = i; cj_[k] = j;
    dk2[k] = dk * dk;
    mi[k] = Sp.invmass[i]; mj[k] = Sp.invmass[j];
    // the factor multiplying (d^2 - r'^2): mode 2 (shake.cpp:67) 0.5 mred_inv,
    // mode 1 (textbook) 0.5 / mred_inv
    hm[k] = R.shake == 2 ? 0.5 * (mi[k] + mj[k]) : 0.5 / (mi[k] + mj[k]);
    for (int q = 0; q < 3; q++) rk[k][q] = r(i, q) - r(j, q);
  }
  int it = 0;
  bool converge = false;
  while (!converge) {
    if (it >= max_iter) return -it;
    it++;
    converge = true;
#pragma unroll
    for (int k = 0; k < kMaxRigid; k++) {
      if (k >= nr) break;
      const int i = ci_[k], j = cj_[k];
      double rn[3];
      for (int q = 0; q < 3; q++) rn[q] = x(i, q) - x(j, q);
      double rn2 = rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2];
      double dot = rk[k][0] * rn[0] + rk[k][1] * rn[1] + rk[k][2] * rn[2];
      if (R.shake != 2) {  // mode 1 tests each constraint before its update
        double d0 = rn2 - dk2[k];
        if (d0 * d0 > tolerance) converge = false;
      }
      double mred_inv = mi[k] + mj[k];
      // mode 2: shake.cpp:67 multiplies by the inverse reduced mass; mode 1
      // divides by it (textbook SHAKE); both reach the same fixed point.
      // (x / 2 * m equals x * (0.5 m) exactly, so mode 2 keeps its rounding)
      double ramk = (dk2[k] - rn2) * hm[k];
      if (dot > 1e-10) ramk /= dot;
      else {
        double rk2 = rk[k][0] * rk[k][0] + rk[k][1] * rk[k][1] + rk[k][2] * rk[k][2];
        ramk = 0.5 * rho_zero + ramk / (rk2 * rho_zero) * mred_inv;
      }
      double cfi = ramk * mi[k], cfj = ramk * mj[k];
      for (int q = 0; q < 3; q++) {
        double ni = x(i, q) + rk[k][q] * cfi, nj = x(j, q) - rk[k][q] * cfj;
        x(i, q) = ni;
        x(j, q) = nj;
        rn[q] = ni - nj;
      }
      rn2 = rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2];
      double dr = rn2 - dk2[k];
      if (!(dr * dr <= 1e30)) return -it;  // diverged (or non-finite)
      if (R.shake == 2 && dr * dr > tolerance) converge = false;
    }
  }
  return it;
}

// Constraint tables of the two constrained species (ion 0, vapor 1), staged
// in LDS once per launch: [2][cap] endpoints, d^2, inverse masses and the
// factor multiplying (d^2 - r'^2) (mode 2: 0.5 mred_inv, shake.cpp:67;
// mode 1: 0.5 / mred_inv), and each species' constrained atoms (bit a of
// cons, molecules of up to 64 atoms).
struct ShakeLds {
  int *n;        // [2]
  int *ij;       // [2][cap][2]
  double *c;     // [2][cap][4]: d^2, 1/m_i, 1/m_j, factor
  unsigned long long *cons;  // [2]
  int *ng, *gs;              // [2], [2][5]: constraint groups (DevSpecies)
  unsigned long long *gm;    // [2][4]
  int *star;                 // [2]: every group a star of <= kStarMax constraints
  int cap;
};

// every constraint group of the species joins one centre atom to distinct
// satellites, at most kStarMax of them (X-H, X-H2, X-H3): DevSpecies::shake_star
constexpr int kStarMax = 3;

__device__ inline void shake_lds_fill(const DevReplica &R, const DevSpecies &Sp, const ShakeLds &T, int s) {
  const int n = Sp.nrigid < T.cap ? Sp.nrigid : T.cap;
  unsigned long long cons = 0;
  for (int k = 0; k < n; k++) {
    int i = Sp.rigid[3 * k], j = Sp.rigid[3 * k + 1];
    double dk = R.btype[2 * Sp.rigid[3 * k + 2] + 1];
    double mi = Sp.invmass[i], mj = Sp.invmass[j];
    int *ij = T.ij + 2 * ((size_t)s * T.cap + k);
    double *c = T.c + 4 * ((size_t)s * T.cap + k);
    ij[0] = i; ij[1] = j;
    c[0] = dk * dk; c[1] = mi; c[2] = mj;
    c[3] = R.shake == 2 ? 0.5 * (mi + mj) : 0.5 / (mi + mj);
    if (i < 64) cons |= 1ull << i;
    if (j < 64) cons |= 1ull << j;
  }
  T.n[s] = Sp.nrigid;
  T.cons[s] = cons;
  T.ng[s] = Sp.shake_ng;
  T.star[s] = Sp.nrigid <= T.cap && Sp.shake_star ? 1 : 0;
  for (int g = 0; g < 5; g++) T.gs[5 * s + g] = Sp.shake_gs[g];
  for (int g = 0; g < 4; g++) T.gm[4 * s + g] = Sp.shake_gm[g];
}

// shake_molecule with the constraint data (and the old bond vectors) read
// from LDS in every sweep: one code shape for any constraint count
template <class XA, class RA>
__device__ int shake_molecule_lds(const DevReplica &R, const ShakeLds &T, int s, XA x, RA r, int max_iter) {
  const int nr = T.n[s];
  const int *ij = T.ij + 2 * (size_t)s * T.cap;
  const double *cc = T.c + 4 * (size_t)s * T.cap;
  const double tolerance = 1e-20, rho_zero = 0.01;
  int it = 0;
  bool converge = false;
  while (!converge) {
    if (it >= max_iter) return -it;
    it++;
    converge = true;
    for (int k = 0; k < nr; k++) {
      const int i = ij[2 * k], j = ij[2 * k + 1];
      const double dk2 = cc[4 * k], mi = cc[4 * k + 1], mj = cc[4 * k + 2], hm = cc[4 * k + 3];
      double rk[3], rn[3];
      for (int q = 0; q < 3; q++) { rk[q] = r(i, q) - r(j, q); rn[q] = x(i, q) - x(j, q); }
      double rn2 = rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2];
      double dot = rk[0] * rn[0] + rk[1] * rn[1] + rk[2] * rn[2];
      if (R.shake != 2) {
        double d0 = rn2 - dk2;
        if (d0 * d0 > tolerance) converge = false;
      }
      double ramk = (dk2 - rn2) * hm;
      // 1 / dot by the hardware reciprocal estimate refined with two
      // Newton-Raphson steps (r <- r + r (1 - dot r)): within an ulp of the
      // quotient, without the dependent chain of a correctly rounded divide
      if (dot > 1e-10) ramk *= rcp_nr(dot);
      else {
        double rk2 = rk[0] * rk[0] + rk[1] * rk[1] + rk[2] * rk[2];
        ramk = 0.5 * rho_zero + ramk / (rk2 * rho_zero) * (mi + mj);
      }
      double cfi = ramk * mi, cfj = ramk * mj;
      for (int q = 0; q < 3; q++) {
        double ni = x(i, q) + rk[q] * cfi, nj = x(j, q) - rk[q] * cfj;
        x(i, q) = ni;
        x(j, q) = nj;
        rn[q] = ni - nj;
      }
      rn2 = rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2];
      double dr = rn2 - dk2;
      if (!(dr * dr <= 1e30)) return -it;
      if (R.shake == 2 && dr * dr > tolerance) converge = false;
    }
  }
  return it;
}

// shake_molecule_lds for one constraint group g of a molecule whose G
// groups (1, 2 or 4) sit on adjacent lanes of one quad: each lane sweeps its
// own group, and after every sweep the molecule's lanes agree on whether all
// groups converged (one more sweep for all while any has not, as in the
// single loop over every constraint) or one ran away.  The groups move
// disjoint atoms, so the positions are those of shake_molecule_lds, to the
// bit; a sweep takes as long as the largest group.
__device__ inline int quad_and(int v, int G) {
  if (G >= 2) v &= __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xf, 0xf, false);  // lanes 1 0 3 2
  if (G >= 4) v &= __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xf, 0xf, false);  // lanes 2 3 0 1
  return v;
}
// The sweep of a star group (every constraint joins one centre atom to its
// own satellite, up to kStarMax: X-H, X-H2, X-H3) with the group's atoms
// held in registers for the whole iteration and written back once.  Each
// constraint is taken as (centre, satellite) whichever order the file gives:
// that only negates the old and new bond vectors, exactly, so every update
// rounds as in the (i, j) order.  Sweep count (negative on failure) as
// shake_group_lds.  For species whose groups are all stars (ShakeLds::star),
// so that all lanes of a molecule take the same path.
template <class XA, class RA>
__device__ int shake_group_star(const DevReplica &R, const ShakeLds &T, int s, int k0, int k1, int G, XA x, RA r,
                                int max_iter) {
  const int n = k1 - k0;
  const int *ij = T.ij + 2 * (size_t)s * T.cap;
  const double *cc = T.c + 4 * (size_t)s * T.cap;
  const int i0 = ij[2 * k0], j0 = ij[2 * k0 + 1];
  const int c = n >= 2 && !(i0 == ij[2 * k0 + 2] || i0 == ij[2 * k0 + 3]) ? j0 : i0;
  const double tolerance = 1e-20, rho_zero = 0.01;
  // the positions, which carry the chain from one update to the next, and
  // every constraint's fixed data (old bond vector, squared length, masses)
  // sit in registers for the whole iteration: a sweep reads no LDS
  double xc[3], xs[kStarMax][3], rk[kStarMax][3], dk2[kStarMax], hm[kStarMax], mc[kStarMax], ms[kStarMax];
  for (int q = 0; q < 3; q++) xc[q] = x(c, q);
#pragma unroll
  for (int k = 0; k < kStarMax; k++) {
    const int kk = k0 + (k < n ? k : 0);
    const bool ci = ij[2 * kk] == c;
    const int sk = ci ? ij[2 * kk + 1] : ij[2 * kk];
    for (int q = 0; q < 3; q++) {
      xs[k][q] = x(sk, q);
      rk[k][q] = r(c, q) - r(sk, q);
    }
    dk2[k] = cc[4 * kk];
    hm[k] = cc[4 * kk + 3];
    mc[k] = ci ? cc[4 * kk + 1] : cc[4 * kk + 2];
    ms[k] = ci ? cc[4 * kk + 2] : cc[4 * kk + 1];
  }
  int it = 0;
  bool converge = false;
  while (!converge) {
    if (it >= max_iter) return -it;
    it++;
    int conv = 1, ok = 1;
#pragma unroll
    for (int k = 0; k < kStarMax; k++) {
      if (k >= n) break;
      double rn[3];
      for (int q = 0; q < 3; q++) rn[q] = xc[q] - xs[k][q];
      double rn2 = rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2];
      double dot = rk[k][0] * rn[0] + rk[k][1] * rn[1] + rk[k][2] * rn[2];
      if (R.shake != 2) {
        double d0 = rn2 - dk2[k];
        if (d0 * d0 > tolerance) conv = 0;
      }
      double ramk = (dk2[k] - rn2) * hm[k];
      if (dot > 1e-10) ramk *= rcp_nr(dot);
      else {
        double rk2 = rk[k][0] * rk[k][0] + rk[k][1] * rk[k][1] + rk[k][2] * rk[k][2];
        ramk = 0.5 * rho_zero + ramk / (rk2 * rho_zero) * (mc[k] + ms[k]);
      }
      double cfc = ramk * mc[k], cfs = ramk * ms[k];
      for (int q = 0; q < 3; q++) {
        double nc = xc[q] + rk[k][q] * cfc, ns = xs[k][q] - rk[k][q] * cfs;
        xc[q] = nc;
        xs[k][q] = ns;
        rn[q] = nc - ns;
      }
      rn2 = rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2];
      double dr = rn2 - dk2[k];
      if (!(dr * dr <= 1e30)) { ok = 0; break; }
      if (R.shake == 2 && dr * dr > tolerance) conv = 0;
    }
    if (!quad_and(ok, G)) return -it;
    converge = quad_and(conv, G) != 0;
  }
  for (int q = 0; q < 3; q++) x(c, q) = xc[q];
#pragma unroll
  for (int k = 0; k < kStarMax; k++) {
    if (k >= n) break;
    const int kk = k0 + k;
    const int sk = ij[2 * kk] == c ? ij[2 * kk + 1] : ij[2 * kk];
    for (int q = 0; q < 3; q++) x(sk, q) = xs[k][q];
  }
  return it;
}

template <class XA, class RA>
__device__ int shake_group_lds(const DevReplica &R, const ShakeLds &T, int s, int g, int G, XA x, RA r,
                               int max_iter) {
  const int k0 = T.gs[5 * s + g], k1 = T.gs[5 * s + g + 1];
  if (T.star[s]) return shake_group_star(R, T, s, k0, k1, G, x, r, max_iter);
  const int *ij = T.ij + 2 * (size_t)s * T.cap;
  const double *cc = T.c + 4 * (size_t)s * T.cap;
  const double tolerance = 1e-20, rho_zero = 0.01;
  int it = 0;
  bool converge = false;
  while (!converge) {
    if (it >= max_iter) return -it;
    it++;
    int conv = 1, ok = 1;
    for (int k = k0; k < k1; k++) {
      const int i = ij[2 * k], j = ij[2 * k + 1];
      const double dk2 = cc[4 * k], mi = cc[4 * k + 1], mj = cc[4 * k + 2], hm = cc[4 * k + 3];
      double rk[3], rn[3];
      for (int q = 0; q < 3; q++) { rk[q] = r(i, q) - r(j, q); rn[q] = x(i, q) - x(j, q); }
      double rn2 = rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2];
      double dot = rk[0] * rn[0] + rk[1] * rn[1] + rk[2] * rn[2];
      if (R.shake != 2) {
        double d0 = rn2 - dk2;
        if (d0 * d0 > tolerance) conv = 0;
      }
      double ramk = (dk2 - rn2) * hm;
      if (dot > 1e-10) ramk *= rcp_nr(dot);
      else {
        double rk2 = rk[0] * rk[0] + rk[1] * rk[1] + rk[2] * rk[2];
        ramk = 0.5 * rho_zero + ramk / (rk2 * rho_zero) * (mi + mj);
      }
      double cfi = ramk * mi, cfj = ramk * mj;
      for (int q = 0; q < 3; q++) {
        double ni = x(i, q) + rk[q] * cfi, nj = x(j, q) - rk[q] * cfj;
        x(i, q) = ni;
        x(j, q) = nj;
        rn[q] = ni - nj;
      }
      rn2 = rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2];
      double dr = rn2 - dk2;
      if (!(dr * dr <= 1e30)) { ok = 0; break; }
      if (R.shake == 2 && dr * dr > tolerance) conv = 0;
    }
    if (!quad_and(ok, G)) return -it;
    converge = quad_and(conv, G) != 0;
  }
  return it;
}

// shake_velocity over the atoms the table marks, (x - x_old) * (1 / dt)
// (the same as the quotient when dt is a power of two)
template <class XA, class RA, class VA>
__device__ void shake_velocity_lds(const ShakeLds &T, int s, int natoms, XA x, RA r, VA v, double dt) {
  const double inv_dt = 1.0 / dt;
  const unsigned long long cons = T.cons[s];
  const int nr = T.n[s];
  const int *ij = T.ij + 2 * (size_t)s * T.cap;
  for (int a = 0; a < natoms; a++) {
    bool c;
    if (a < 64) c = (cons >> a) & 1ull;
    else {
      c = false;
      for (int k = 0; k < nr && !c; k++) c = ij[2 * k] == a || ij[2 * k + 1] == a;
    }
    if (!c) continue;
    for (int q = 0; q < 3; q++) v(a, q) = (x(a, q) - r(a, q)) * inv_dt;
  }
}

// a failed SHAKE leaves the molecule unconstrained for this step (the
// reference, which never calls shake(), always does): x = x_old + v dt
template <class XA, class RA, class VA>
__device__ void shake_undo(const DevSpecies &Sp, XA x, RA r, VA v, double dt) {
  for (int a = 0; a < Sp.natoms; a++)
    for (int q = 0; q < 3; q++) x(a, q) = r(a, q) + v(a, q) * dt;
}

// SHAKE mode 1: every atom that takes part in a constraint gets the
// constrained half-step velocity (x_shaken - x_old) / dt
template <class XA, class RA, class VA>
__device__ void shake_velocity(const DevSpecies &Sp, XA x, RA r, VA v, double dt) {
  uint64_t cons = 0;  // constrained atoms of molecules with <= 64 atoms
  for (int k = 0; k < Sp.nrigid; k++) {
    int i = Sp.rigid[3 * k], j = Sp.rigid[3 * k + 1];
    if (i < 64) cons |= 1ull << i;
    if (j < 64) cons |= 1ull << j;
  }
  for (int a = 0; a < Sp.natoms; a++) {
    bool c;
    if (a < 64) c = (cons >> a) & 1ull;
    else {
      c = false;
      for (int k = 0; k < Sp.nrigid && !c; k++) c = Sp.rigid[3 * k] == a || Sp.rigid[3 * k + 1] == a;
    }
    if (!c) continue;
    for (int q = 0; q < 3; q++) v(a, q) = (x(a, q) - r(a, q)) / dt;
  }
}

// ------------------------------------------------------------------------
// MD::make_pair (pairlist.cpp:16-49) for one replica, by one workgroup, on
// global memory.  FP decisions (advance, COM, classification, list radii,
// v_max) use the reference's operation order with contraction disabled, so
// that given the same state the lists are bit-identical.  s_int >= 64 ints.
__device__ inline void enter_molecule(const DevReplica &R, int i, int sp, long long rb) {
#pragma clang fp contract(off)
  const DevSpecies &S = R.sp[sp];
  int na = S.natoms;
  int o = sp == 1 ? R.off_gas + (i - 1) * na : R.off_vap + (i - 1 - R.n_gas) * na;
  double *x = R.atom_x + 3 * (size_t)o, *v = R.atom_v + 3 * (size_t)o;
  for (int q = 0; q < 3 * na; q++) { x[q] = S.tmpl[q]; v[q] = 0; }
  if (sp == 2)
    for (int q = 0; q < 3 * na; q++) R.atom_f[3 * (size_t)o + q] = 0;
  if (sp == 1 && R.gas_spin && na >= 2) {
    // N2 internal rotation, updateGasIn.cpp:66-76
    double z[2];
    rng_normal2(R.key0, R.key1, (uint32_t)i, rb, 4, z);
    double vy = R.sig_gas * z[0] * 1e-5;
    v[0] = 0; v[1] = vy; v[2] = vy;
    v[3] = 0; v[4] = -vy; v[5] = -vy;
  }
  double u[2], u3[2];
  rng_uniform2(R.key0, R.key1, (uint32_t)i, rb, 2, u);
  rng_uniform2(R.key0, R.key1, (uint32_t)i, rb, 3, u3);
  double a = u[0] * 2 * kPI, b = u[1] * 2 * kPI, c = u3[0] * 2 * kPI;
  const double *X = R.mol_x + 3 * (size_t)i, *V = R.mol_v + 3 * (size_t)i;
  for (int ii = 0; ii < na; ii++) {
    double *p = x + 3 * ii;
    // the gas variant passes (x, y, y), updateGasIn.cpp:84
    double za = sp == 1 ? p[1] : p[2];
    rotation(p[0], p[1], p[2], a, b, c, p[0], p[1], za);
    p[0] += X[0]; p[1] += X[1]; p[2] += X[2];
  }
  for (int ii = 0; ii < na; ii++) {
    double *p = v + 3 * ii;
    double za = sp == 1 ? p[1] : p[2];
    rotation(p[0], p[1], p[2], a, b, c, p[0], p[1], za);
    p[0] += V[0]; p[1] += V[1]; p[2] += V[2];
  }
}

// s_int: >= 16 ints, s_dbl: >= 16 doubles, s_big: big_cap doubles of
// workgroup scratch (LDS) for the list tests
__device__ inline void rebuild_wg(const DevReplica &R, int *s_int, double *s_dbl, double *s_big, int big_cap) {
#pragma clang fp contract(off)
  const int tid = threadIdx.x, nt = blockDim.x;
  Dyn *D = R.dyn;
  const int lp = D->loopPair;
  const long long rb = D->rebuild_count;
  double *mol = R.mol_x, *molv = R.mol_v;
  const double HL = R.d_size * 0.5;
  const double cx = mol[0], cy = mol[1], cz = mol[2];
  const double cvx = molv[0], cvy = molv[1], cvz = molv[2];
  // advance (out) / COM (in) / classify, one molecule per thread
  for (int i = 1 + tid; i < R.n_mol; i += nt) {
    int sp = i <= R.n_gas ? 1 : 2;
    const DevSpecies &S = R.sp[sp];
    double *m = mol + 3 * (size_t)i, *mv = molv + 3 * (size_t)i;
    bool was_in = R.region[i] != 0;
    if (!was_in) {
      // pairlist.cpp:20-29
      m[0] += mv[0] * R.dt * lp;
      m[1] += mv[1] * R.dt * lp;
      m[2] += mv[2] * R.dt * lp;
    } else {
      // updateCenters.cpp:6-57 (divides by the declared molecule mass)
      int o = sp == 1 ? R.off_gas + (i - 1) * S.natoms : R.off_vap + (i - 1 - R.n_gas) * S.natoms;
      double s0 = 0, s1 = 0, s2 = 0, v0 = 0, v1 = 0, v2 = 0;
      for (int ii = 0; ii < S.natoms; ii++) {
        const double *xa = R.atom_x + 3 * (size_t)(o + ii), *va = R.atom_v + 3 * (size_t)(o + ii);
        double ma = S.mass[ii];
        s0 += xa[0] * ma; s1 += xa[1] * ma; s2 += xa[2] * ma;
        v0 += va[0] * ma; v1 += va[1] * ma; v2 += va[2] * ma;
      }
      double mi = 1 / S.molmass;
      m[0] = s0 * mi; m[1] = s1 * mi; m[2] = s2 * mi;
      mv[0] = v0 * mi; mv[1] = v1 * mi; mv[2] = v2 * mi;
    }
    // updateGasIn.cpp:21-47 / updateVaporIn.cpp:23-51
    double dx = m[0] - cx, dy = m[1] - cy, dz = m[2] - cz;
    double r2 = (dx * dx + dy * dy + dz * dz);
    if (sp == 2 && was_in && R.vlog) {
      // MD::Ovin / MD::Ovout (updateVaporIn.cpp:33, 39), before the wrap and
      // the velocity draw
      int slot = atomicAdd(&D->n_log, 1);
      if (slot < R.log_cap) {
        VapEvent &e = R.vlog[slot];
        e.mol = i;
        e.kind = r2 < kRI2 ? 0 : 1;
        e.step = D->step;
        e.dx[0] = dx; e.dx[1] = dy; e.dx[2] = dz;
        e.dv[0] = mv[0] - cvx; e.dv[1] = mv[1] - cvy; e.dv[2] = mv[2] - cvz;
      }
    }
    if (r2 < kRI2) {
      if (!was_in) enter_molecule(R, i, sp, rb);
      R.region[i] = 1;
    } else {
      R.region[i] = 0;
      if (dx < -HL) m[0] += R.d_size;
      if (dy < -HL) m[1] += R.d_size;
      if (dz < -HL) m[2] += R.d_size;
      if (dx > HL) m[0] -= R.d_size;
      if (dy > HL) m[1] -= R.d_size;
      if (dz > HL) m[2] -= R.d_size;
      double sg = sp == 1 ? R.sig_gas : R.sig_vap;
      // three normals: a pair, then the first of another pair
      double z[2];
      rng_normal2(R.key0, R.key1, (uint32_t)i, rb, 0, z);
      const double z3 = rng_normal1(R.key0, R.key1, (uint32_t)i, rb, 1);
      mv[0] = sg * z[0] * 1e-5;
      mv[1] = sg * z[1] * 1e-5;
      mv[2] = sg * z3 * 1e-5;
    }
  }
  __syncthreads();
  // ordered compaction of the in-domain gas and vapor ids
  const int ng = wave_compact(
      R.n_gas, [&](int k) { return R.region[1 + k] != 0; }, [&](int k, int p) { R.in_gas[p] = 1 + k; }, s_int);
  const int nv = wave_compact(
      R.n_vap, [&](int k) { return R.region[1 + R.n_gas + k] != 0; },
      [&](int k, int p) { R.in_vap[p] = 1 + R.n_gas + k; }, s_int);
  __syncthreads();
  // Potential::makePair lists (GasIon.cpp:41-60, VaporGas.cpp:47-64, gas-gas),
  // ascending gas molecule then partner, stored as in-list positions.  The
  // tests read the ion's atoms and the in-domain COMs from LDS when they fit
  // (global loads would put a memory latency on every test); gi and gv are
  // compacted over their row-major index, so that rows shorter than a wave
  // (a one-atom ion, a few vapor molecules) still fill its lanes.
  const int n_ion = R.sp[0].natoms;
  int n_gi = 0, n_gv = 0, n_gg = 0;
  auto lists = [&](auto ionp, auto gasp, auto vapp) {
    auto near = [&](const double *g, const double *h) {
#pragma clang fp contract(off)
      double dx = g[0] - h[0], dy = g[1] - h[1], dz = g[2] - h[2];
      return (dx * dx + dy * dy + dz * dz) < R.ML2;
    };
    if ((R.flags & B_GI) && n_ion > 0)
      n_gi = wave_compact(
          ng * n_ion, [&](int f) { const int k = f / n_ion; return near(gasp(k), ionp(f - k * n_ion)); },
          [&](int f, int p) {
            const int k = f / n_ion;
            if (p < R.cap_gi) { R.gi[2 * p] = k; R.gi[2 * p + 1] = f - k * n_ion; }
          },
          s_int);
    if ((R.flags & B_GV) && nv > 0)
      n_gv = wave_compact(
          ng * nv, [&](int f) { const int k = f / nv; return near(gasp(k), vapp(f - k * nv)); },
          [&](int f, int p) {
            const int k = f / nv;
            if (p < R.cap_gv) { R.gv[2 * p] = k; R.gv[2 * p + 1] = f - k * nv; }
          },
          s_int);
    if (R.flags & B_GG)
      n_gg = wave_compact_pairs(
          ng, [](int k) { return k + 1; }, [&](int) { return ng; }, [&](int k, int l) { return near(gasp(k), gasp(l)); },
          [&](int k, int l, int p) {
            if (p < R.cap_gg) { R.gg[2 * p] = k; R.gg[2 * p + 1] = l; }
          },
          s_int);
  };
  if (3 * (n_ion + ng + nv) <= big_cap) {
    double *si = s_big, *sg = s_big + 3 * n_ion, *sv = sg + 3 * ng;
    for (int t = tid; t < 3 * n_ion; t += nt) si[t] = R.atom_x[t];
    for (int t = tid; t < 3 * ng; t += nt) sg[t] = mol[3 * (size_t)R.in_gas[t / 3] + t % 3];
    for (int t = tid; t < 3 * nv; t += nt) sv[t] = mol[3 * (size_t)R.in_vap[t / 3] + t % 3];
    __syncthreads();
    lists([&](int j) { return si + 3 * j; }, [&](int k) { return sg + 3 * k; }, [&](int l) { return sv + 3 * l; });
  } else {
    lists([&](int j) { return R.atom_x + 3 * (size_t)j; }, [&](int k) { return mol + 3 * (size_t)R.in_gas[k]; },
          [&](int l) { return mol + 3 * (size_t)R.in_vap[l]; });
  }
  // v_max over every molecule (pairlist.cpp:38-46); max is order-free
  double vm = 0.0;
  for (int i = tid; i < R.n_mol; i += nt) {
    const double *v = molv + 3 * (size_t)i;
    double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (vm < v2) vm = v2;
  }
  for (int o = 32; o > 0; o >>= 1) {
    double y = __shfl_xor(vm, o);
    vm = vm < y ? y : vm;
  }
  if ((tid & 63) == 0) s_dbl[tid >> 6] = vm;
  __syncthreads();
  if (tid == 0) {
    double m2 = 0.0;
    for (int w = 0; w < (nt + 63) / 64; w++) m2 = m2 < s_dbl[w] ? s_dbl[w] : m2;
    double lu = R.margin / (sqrt(m2) * R.dt);
    D->loop_update = lu >= 2147483647.0 ? 2147483647 : (int)lu;
    D->loopPair = 0;
    D->rebuild_count = rb + 1;
    D->n_in_gas = ng;
    D->n_in_vap = nv;
    int st = 0;
    if (n_gi > R.cap_gi) { st |= ST_CAP_GI; n_gi = R.cap_gi; }
    if (n_gv > R.cap_gv) { st |= ST_CAP_GV; n_gv = R.cap_gv; }
    if (n_gg > R.cap_gg) { st |= ST_CAP_GG; n_gg = R.cap_gg; }
    D->n_gi = n_gi; D->n_gv = n_gv; D->n_gg = n_gg;
    D->status |= st;
  }
  __syncthreads();
}

}  // namespace mdn
