// mdn_device.h -- device-side data model and arithmetic of libmdn (gfx950).
//
// One replica = one effective-domain system (ion + carrier gas + vapor) of
// the reference's Variables (src/variables.hpp:6-101), flattened to
// structure-of-arrays device buffers.  Every per-step operation is written
// once here, over a "store" policy that says where the in-domain atoms live:
//   GStore  global memory, one replica per launch (the per-call drop-in path)
//   SStore  LDS, one workgroup per replica (the persistent batched path)
// Both index in-domain atoms in the same local order:
//   [ion atoms | atoms of in-domain vapor molecules | atoms of in-domain gas
//    molecules], molecules in ascending id order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdn {

// constants.hpp:44-48 (effective domain), verlet.cpp:49 (kick factor)
constexpr double kRinter = 100.0, kOverlap = 10.0;
constexpr double kRcenter = kRinter - kOverlap * 0.5;
constexpr double kRi2 = kRinter * kRinter;
constexpr double kRI2 = (kRinter + kOverlap) * (kRinter + kOverlap);
constexpr double kPI = 3.14159265358979323846;
constexpr double kSqrt24 = 4.89897948556635619639;  // sqrt(24)
constexpr double kKickUnit = 4.184e-4;
// constants.hpp:30, 49-50 and observer.hpp:6: Boltzmann constant in kcal/mol/K,
// g/mol A^2/fs^2 -> kcal/mol, and Tion = K / N * (1 / kB) / 1.5
constexpr double kKbReal = 1.38e-23 * 6.02e23 / 4184.0;
constexpr double kRealToKcal = 10000 / 4.184;
// thermostat modes of the ion (NVTion line, setCondition.cpp:158-172)
enum { THERMO_NVE = 0, THERMO_SCALE = 1, THERMO_NH = 2 };

// interaction bits (same values as mdn.h)
enum : uint32_t {
  B_GI = 1u << 0, B_VI = 1u << 1, B_GG = 1u << 2, B_GV = 1u << 3, B_VV = 1u << 4,
  B_AMBER = 1u << 5, B_BMH = 1u << 6, B_TERSOFF = 1u << 7, B_SW = 1u << 8,
  B_EFIELD = 1u << 9, B_VINTRA = 1u << 10, B_GINTRA = 1u << 11
};
enum { U_ION = 0, U_GAS, U_VAP, U_GI, U_GG, U_VG, U_VI, U_VV, U_N };

// status bits raised by kernels (host turns them into mdn_status codes)
enum : int { ST_CAP_GI = 1, ST_CAP_GV = 2, ST_CAP_GG = 4, ST_SHAKE = 8, ST_LDS_CAP = 16 };

struct DevSpecies {
  int natoms, nbond, nangle, ndih, nrigid;
  double molmass;
  const double *mass, *invmass, *charge, *tmpl;
  const double *se, *sg;  // per atom sqrt(eps) and sigma of its type
  const int *type, *bond, *angle, *dih, *rigid;
  // the constraints in groups that share no atom (at most 4; one when the
  // molecule has more than 64 atoms): group g is rigid triples
  // [shake_gs[g], shake_gs[g + 1]) and moves the atoms of bit mask
  // shake_gm[g].  Within a group the constraints keep their file order.
  int shake_ng, shake_gs[5];
  unsigned long long shake_gm[4];
  // 1 when every group joins one centre atom to distinct satellites, at
  // most three (X-H, X-H2, X-H3): SHAKE then iterates it in registers
  int shake_star;
};

// mutable per-replica scalars
struct Dyn {
  double U[U_N];
  long long rebuild_count, atom_steps;
  int loopPair, loop_update;
  int n_gi, n_gv, n_gg;
  int n_in_gas, n_in_vap;
  int status, shake_iters_max;
  int thermo;                // THERMO_* of the ion
  // 1: the batched kernel stopped this replica inside a step (its in-domain
  // atoms outgrew the launch's LDS at a rebuild): positions, weights and
  // lists are those of the step, velocities have their first half-kick, the
  // forces and the second half-kick are still to come
  int pending;
  unsigned long long shake_fail;
  double zeta_ion, Tnh_ion;  // Nose-Hoover friction and target (nvt.cpp:19-35)
  long long step;            // MD::itime: steps since the caller last set it
  int n_log, pad2;           // vapor events recorded since the last drain
};

// one Ovin / Ovout record (output.cpp:43-64) of a rebuild: a vapor molecule
// that was in the domain and stays in (kind 0) or leaves (kind 1), relative
// to the ion COM; same layout as mdn_vapor_event
struct VapEvent {
  int mol, kind;
  long long step;
  double dx[3], dv[3];
};

struct DevReplica {
  DevSpecies sp[3];
  int n_gas, n_vap, n_mol, n_atoms, n_types;
  int off_gas, off_vap;
  uint32_t flags;
  int shake, shake_max_iter, gas_spin, ion_type_base;
  double dt, d_size, margin, ML2, qqrd2e, sig_gas, sig_vap;
  double efield[3];
  uint32_t key0, key1;
  const double *pair_coeff, *btype, *ctype, *dcoeff;
  const int *dmulti;
  int amber_words;              // words per row of amber_allow
  const uint32_t *amber_allow;  // [n_ion][amber_words]: bit j of row i = pair (i,j) not excluded
  double *mol_x, *mol_v;
  uint8_t *region;
  double *atom_x, *atom_v, *atom_f, *atom_w, *atom_xo;
  int *in_gas, *in_vap;
  int *gi, *gv, *gg;  // local-index lists (in-list positions), 2 ints per entry
  int cap_gi, cap_gv, cap_gg;
  Dyn *dyn;
  VapEvent *vlog;  // [log_cap] or null (vapor events not recorded)
  int log_cap;
};

// ------------------------------------------------------------------------
// Random stream: Philox4x32-10 (Salmon et al., SC'11), keyed by (seed,
// replica) and countered by (molecule, rebuild, purpose).  The same stream is
// specified for the host, so any host C++ caller can reproduce every draw.
__host__ __device__ inline void philox4x32_10(const uint32_t ci[4], const uint32_t ki[2], uint32_t out[4]) {
  uint32_t c0 = ci[0], c1 = ci[1], c2 = ci[2], c3 = ci[3], k0 = ki[0], k1 = ki[1];
  for (int r = 0; r < 10; r++) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

__host__ __device__ inline void rng_words(uint32_t k0, uint32_t k1, uint32_t mol, long long rebuild, uint32_t purpose,
                                          uint32_t w[4]) {
  uint32_t ctr[4] = {mol, (uint32_t)(unsigned long long)rebuild, (uint32_t)((unsigned long long)rebuild >> 32),
                     purpose};
  uint32_t key[2] = {k0, k1};
  philox4x32_10(ctr, key, w);
}
__host__ __device__ inline double u01_53(uint32_t hi, uint32_t lo) {
  uint64_t v = ((uint64_t)hi << 32) | lo;
  return (double)(v >> 11) * (1.0 / 9007199254740992.0);
}
__host__ __device__ inline void rng_uniform2(uint32_t k0, uint32_t k1, uint32_t mol, long long rb, uint32_t purpose,
                                             double u[2]) {
  uint32_t w[4];
  rng_words(k0, k1, mol, rb, purpose, w);
  u[0] = u01_53(w[0], w[1]);
  u[1] = u01_53(w[2], w[3]);
}
// Box-Muller on (0,1] x [0,1): z = r (cos th, sin th), the sine and cosine
// from one argument reduction
__device__ inline void rng_normal2(uint32_t k0, uint32_t k1, uint32_t mol, long long rb, uint32_t purpose,
                                   double z[2]) {
  uint32_t w[4];
  rng_words(k0, k1, mol, rb, purpose, w);
  uint64_t v = ((uint64_t)w[0] << 32) | w[1];
  double u1 = (double)((v >> 11) + 1) * (1.0 / 9007199254740992.0);
  double u2 = u01_53(w[2], w[3]);
  double r = sqrt(-2.0 * log(u1));
  double th = 2.0 * kPI * u2;
  double sn, cs;
  sincos(th, &sn, &cs);
  z[0] = r * cs;
  z[1] = r * sn;
}
// the first of rng_normal2's pair alone (r cos th)
__device__ inline double rng_normal1(uint32_t k0, uint32_t k1, uint32_t mol, long long rb, uint32_t purpose) {
  uint32_t w[4];
  rng_words(k0, k1, mol, rb, purpose, w);
  uint64_t v = ((uint64_t)w[0] << 32) | w[1];
  double u1 = (double)((v >> 11) + 1) * (1.0 / 9007199254740992.0);
  double u2 = u01_53(w[2], w[3]);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * kPI * u2);
}

// Variables::ROTATION, src/variables.hpp:94-98
__device__ inline void rotation(double &X, double &Y, double &Z, double A, double B, double C, double x, double y,
                                double z) {
  double cA = cos(A), sA = sin(A), cB = cos(B), sB = sin(B), cC = cos(C), sC = sin(C);
  double t = x * cB + y * sA * sB - z * cA * sB;
  double u = y * cA + z * sA;
  X = cC * t + sC * u;
  Y = -sC * t + cC * u;
  Z = x * sB - y * sA * cB + z * cA * cB;
}

// blend weight of an atom at squared distance r2 from the ion COM,
// src/verlet.cpp:117-130 (Rcenter = 95: 1 -> 0.5 at 100 A, 0 at 105 A,
// 0.5 at 110 A^-, 0 beyond)
__device__ inline double blend_weight(double r2) {
  if (r2 < kRi2) return 1.0;
  if (r2 > kRI2) return 0.0;
  double dr = sqrt(r2) - kRcenter;
  double C = cos(kPI / kOverlap * 0.5 * dr);
  return C * C;
}
// the weight in pair precision: FP32 sqrt/cos in mixed mode (the weight only
// scales FP32 pair forces there)
template <typename R>
__device__ inline R blend_weight_r(double r2) { return (R)blend_weight(r2); }
template <>
__device__ inline float blend_weight_r<float>(double r2) {
  if (r2 < kRi2) return 1.0f;
  if (r2 > kRI2) return 0.0f;
  float dr = sqrtf((float)r2) - (float)kRcenter;
  float C = __cosf((float)(kPI / kOverlap * 0.5) * dr);
  return C * C;
}

// ------------------------------------------------------------------------
// Forward-mode dual number: a value and its derivative along one direction
// of the positions.  Instantiating the force routines with it gives the
// derivative of the forces along u, i.e. the Hessian of -U applied to u,
// exactly (to FP64 rounding).  Comparisons look at the value only.
struct Dual {
  double v, d;
  __host__ __device__ Dual() : v(0), d(0) {}
  __host__ __device__ Dual(double x) : v(x), d(0) {}
  __host__ __device__ Dual(double x, double dx) : v(x), d(dx) {}
  __host__ __device__ explicit operator double() const { return v; }
  __device__ Dual &operator+=(const Dual &b) { v += b.v; d += b.d; return *this; }
  __device__ Dual &operator-=(const Dual &b) { v -= b.v; d -= b.d; return *this; }
  __device__ Dual &operator*=(const Dual &b) { d = d * b.v + v * b.d; v *= b.v; return *this; }
};
__device__ inline Dual operator-(const Dual &a) { return Dual(-a.v, -a.d); }
__device__ inline Dual operator+(const Dual &a, const Dual &b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ inline Dual operator-(const Dual &a, const Dual &b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ inline Dual operator*(const Dual &a, const Dual &b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ inline Dual operator/(const Dual &a, const Dual &b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ inline Dual operator+(const Dual &a, double b) { return Dual(a.v + b, a.d); }
__device__ inline Dual operator+(double a, const Dual &b) { return Dual(a + b.v, b.d); }
__device__ inline Dual operator-(const Dual &a, double b) { return Dual(a.v - b, a.d); }
__device__ inline Dual operator-(double a, const Dual &b) { return Dual(a - b.v, -b.d); }
__device__ inline Dual operator*(const Dual &a, double b) { return Dual(a.v * b, a.d * b); }
__device__ inline Dual operator*(double a, const Dual &b) { return Dual(a * b.v, a * b.d); }
__device__ inline Dual operator/(const Dual &a, double b) { return Dual(a.v / b, a.d / b); }
__device__ inline Dual operator/(double a, const Dual &b) {
  const double q = a / b.v;
  return Dual(q, -q * b.d / b.v);
}
__device__ inline bool operator>(const Dual &a, double b) { return a.v > b; }
__device__ inline bool operator<(const Dual &a, double b) { return a.v < b; }
// the overloads below sit beside the usual double ones
using ::acos;
using ::sqrt;
__device__ inline Dual sqrt(const Dual &a) {
  const double r = ::sqrt(a.v);
  return Dual(r, r > 0 ? 0.5 * a.d / r : 0.0);
}
// d acos(x) = -dx / sqrt(1 - x^2)
__device__ inline Dual acos(const Dual &a) { return Dual(::acos(a.v), -a.d / ::sqrt(1.0 - a.v * a.v)); }
__device__ inline double val(double a) { return a; }
__device__ inline double val(const Dual &a) { return a.v; }

// ------------------------------------------------------------------------
// Pair arithmetic.  R = double (plain) or float (mixed: the displacement is
// formed in FP64 and rounded, the rest is FP32; the caller accumulates the
// force in FP64).
// 1/r^2 and 1/r: IEEE in FP64 (the reference's 1/rsq and sqrt(r2inv)); in
// FP32 the hardware reciprocal square root (v_rsq_f32, about 1 ulp)
template <typename R>
__device__ inline void inv_r(R rsq, R &r2inv, R &rinv) {
  r2inv = R(1) / rsq;
  rinv = sqrt(r2inv);
}
template <>
__device__ inline void inv_r<float>(float rsq, float &r2inv, float &rinv) {
  rinv = __builtin_amdgcn_rsqf(rsq);
  r2inv = rinv * rinv;
}
template <typename R>
__device__ inline R inv_r2(R rsq) { return R(1) / rsq; }
template <>
__device__ inline float inv_r2<float>(float rsq) { return __builtin_amdgcn_rcpf(rsq); }

// LJ + Coulomb, VaporIon.cpp:19-37 / VaporVapor.cpp:21-39 / computeLong.
template <typename R>
__device__ inline R ljcoul_fpair(R rsq, R c0, R c1, R qq, R w, bool eflag, double &e) {
  R r2inv, rinv;
  inv_r<R>(rsq, r2inv, rinv);
  R r6inv = r2inv * r2inv * r2inv;
  R flj = r6inv * (c0 * r6inv - c1);
  R fc = qq * rinv;
  if (eflag) e += (double)(r6inv * (c0 / R(12) * r6inv - c1 / R(6))) + (double)fc;
  return (flj + fc) * r2inv * w;
}
// LJ only, GasIon.cpp:19-33 / VaporGas.cpp:19-35
template <typename R>
__device__ inline R lj_fpair(R rsq, R c0, R c1, R w, bool eflag, double &e) {
  R r2inv = inv_r2<R>(rsq);
  R r6inv = r2inv * r2inv * r2inv;
  R flj = r6inv * (c0 * r6inv - c1);
  if (eflag) e += (double)(r6inv * (c0 / R(12) * r6inv - c1 / R(6)));
  return flj * r2inv * w;
}

// harmonic bond, compute_AMBER.cpp:52-77: returns fb with F_i = fb*d, F_j = -fb*d
// (T = double, or Dual for the derivative along a direction)
template <typename T>
__device__ inline T bond_fb(T rsq, double k, double r0, bool eflag, double &e) {
  T r = sqrt(rsq);
  T dr = r - r0;
  T rk = k * dr;
  if (eflag) e += val(rk * dr);
  return -2.0 * rk / r;
}

// harmonic angle, compute_AMBER.cpp:80-127
template <typename T>
__device__ inline void angle_f(const T d1[3], const T d2[3], double k, double th0, bool eflag, double &e, T f1[3],
                               T f3[3]) {
  T rsq1 = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
  T rsq2 = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
  T r1 = sqrt(rsq1), r2 = sqrt(rsq2);
  T C = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
  C = C / (r1 * r2);
  T Cs = 1.0 / sqrt(1.0 - C * C);
  T dtheta = acos(C) - th0;
  T tk = k * dtheta;
  if (eflag) e += val(tk * dtheta);
  T a = -2.0 * tk * Cs;
  T a11 = a * C / rsq1, a12 = -a / (r1 * r2), a22 = a * C / rsq2;
  for (int q = 0; q < 3; q++) {
    f1[q] = a11 * d1[q] + a12 * d2[q];
    f3[q] = a22 * d2[q] + a12 * d1[q];
  }
}

// multi-term dihedral, compute_AMBER.cpp:130-253 (LAMMPS charmm form, n-fold
// recurrence for cos/sin(n phi); the n==0 special case checks the first
// term's n for every term, compute_AMBER.cpp:199)
template <typename T>
__device__ inline void dihedral_f(const T vb1[3], const T vb2[3], const T vb3[3], const double *co, int multi,
                                  bool eflag, double &e, T f1[3], T f2[3], T f3[3], T f4[3]) {
  T vb2m[3] = {-vb2[0], -vb2[1], -vb2[2]};
  T ax = vb1[1] * vb2m[2] - vb1[2] * vb2m[1];
  T ay = vb1[2] * vb2m[0] - vb1[0] * vb2m[2];
  T az = vb1[0] * vb2m[1] - vb1[1] * vb2m[0];
  T bx = vb3[1] * vb2m[2] - vb3[2] * vb2m[1];
  T by = vb3[2] * vb2m[0] - vb3[0] * vb2m[2];
  T bz = vb3[0] * vb2m[1] - vb3[1] * vb2m[0];
  T rasq = ax * ax + ay * ay + az * az;
  T rbsq = bx * bx + by * by + bz * bz;
  T rgsq = vb2m[0] * vb2m[0] + vb2m[1] * vb2m[1] + vb2m[2] * vb2m[2];
  T rg = sqrt(rgsq);
  T rginv = rg > 0 ? 1.0 / rg : T(0.0);
  T ra2inv = rasq > 0 ? 1.0 / rasq : T(0.0);
  T rb2inv = rbsq > 0 ? 1.0 / rbsq : T(0.0);
  T rabinv = sqrt(ra2inv * rb2inv);
  T c = (ax * bx + ay * by + az * bz) * rabinv;
  T s = rg * rabinv * (ax * vb3[0] + ay * vb3[1] + az * vb3[2]);
  T df = 0.0;
  for (int JJ = 0; JJ < multi; JJ++) {
    const double *t = co + 5 * JJ;
    T p = 1.0, df1 = 0.0, ddf1 = 0.0;
    for (int l = 0; l < t[1]; l++) {
      ddf1 = p * c - df1 * s;
      df1 = p * s + df1 * c;
      p = ddf1;
    }
    p = p * t[3] + df1 * t[4];
    df1 = df1 * t[3] - ddf1 * t[4];
    df1 *= -t[1];
    p += 1.0;
    if (co[1] == 0) {
      p = 1.0 + t[3];
      df1 = 0.0;
    }
    df += -t[0] * df1;
    if (eflag) e += val(t[0] * p);
  }
  T fg = vb1[0] * vb2m[0] + vb1[1] * vb2m[1] + vb1[2] * vb2m[2];
  T hg = vb3[0] * vb2m[0] + vb3[1] * vb2m[1] + vb3[2] * vb2m[2];
  T fga = fg * ra2inv * rginv, hgb = hg * rb2inv * rginv;
  T gaa = -ra2inv * rg, gbb = rb2inv * rg;
  T dtf[3] = {gaa * ax, gaa * ay, gaa * az};
  T dtg[3] = {fga * ax - hgb * bx, fga * ay - hgb * by, fga * az - hgb * bz};
  T dth[3] = {gbb * bx, gbb * by, gbb * bz};
  for (int q = 0; q < 3; q++) {
    T sq = df * dtg[q];
    f1[q] = df * dtf[q];
    f2[q] = sq - f1[q];
    f4[q] = df * dth[q];
    f3[q] = -sq - f4[q];
  }
}

// ------------------------------------------------------------------------
// Workgroup helpers (nthr = blockDim.x, multiple of 64, at most 1024)
// 1 / d: the hardware reciprocal estimate refined by two Newton-Raphson
// steps, r <- r + r (1 - d r) (each roughly doubles the correct bits)
__device__ inline double rcp_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  r = fma(r, fma(-d, r, 1.0), r);
  return r;
}
__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
// a DPP move of a double: lanes whose source is outside the control's
// reach (or in a row the row mask leaves out) read 0
template <int CTRL, int ROWS>
__device__ inline double dpp_take(double v) {
  const long long b = __double_as_longlong(v);
  int lo, hi;
  if (ROWS == 0xf) {  // every row: bound_ctrl gives the 0
    lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xf, 0xf, true);
    hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, true);
  } else {  // rows left out keep the old value, 0
    lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWS, 0xf, false);
    hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, false);
  }
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// wave_sum for a wave whose 64 lanes are all active, without the LDS
// permute path: inclusive sums within each row of 16 lanes (row_shr 1, 2,
// 4, 8), then row 0's total into row 1 and row 2's into row 3
// (row_bcast:15), rows 0-1 into rows 2-3 (row_bcast:31); lane 63 then
// holds the total, which every lane reads
__device__ inline double wave_sum_full(double v) {
  v += dpp_take<0x111, 0xf>(v);
  v += dpp_take<0x112, 0xf>(v);
  v += dpp_take<0x114, 0xf>(v);
  v += dpp_take<0x118, 0xf>(v);
  v += dpp_take<0x142, 0xa>(v);
  v += dpp_take<0x143, 0xc>(v);
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// Tion of Observer::computeProps (observer.cpp:27-36) from sum m v^2 over
// the N ion atoms
__host__ __device__ inline double ion_temperature(double mv2, int n) {
  return mv2 * (0.5 * kRealToKcal) / (double)n * ((1 / kKbReal) / 1.5);
}
// MD::nosehoover_zeta, nvt.cpp:18-24 (Q^-1 = 1e-4, g = 3N)
__host__ __device__ inline double nh_zeta_step(double zeta, double Tion, double Tnh, int n, double dt) {
  return zeta + (Tion - Tnh) * (n * 3) * kKbReal * 0.0001 * dt;
}
// MD::nosehoover_ion, nvt.cpp:27-35
__host__ __device__ inline double nh_factor(double zeta, double dt) { return exp(-zeta * 0.5 * dt); }
// exclusive scan of one int per thread over the workgroup; s_tmp holds
// >= nthr/64 ints.  Returns the prefix, *total = the workgroup sum.
__device__ inline int block_scan(int v, int *s_tmp, int *total) {
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_tmp[wave] = x;
  __syncthreads();
  int base = 0, tot = 0;
  for (int w = 0; w < nw; w++) {
    int c = s_tmp[w];
    if (w < wave) base += c;
    tot += c;
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

// Ordered compaction over the workgroup without a scan per chunk: each wave
// takes a contiguous range [n * w / nw, n * (w + 1) / nw) of the items,
// counts its hits with ballots, and after one exchange of the per-wave
// counts (s_tmp: >= nthr/64 ints) writes them in item order.  keep(i)
// tests item i; put(i, p) stores it at output position base + p.  Returns
// the workgroup's hit count (the same on every thread).
template <class Keep, class Put>
__device__ inline int wave_compact(int n, Keep keep, Put put, int *s_tmp) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  const int i0 = (int)((long long)n * wave / nw), i1 = (int)((long long)n * (wave + 1) / nw);
  int cnt = 0;
  for (int c = i0; c < i1; c += 64) {
    const int i = c + lane;
    cnt += __popcll(__ballot(i < i1 && keep(i)));
  }
  if (lane == 0) s_tmp[wave] = cnt;
  __syncthreads();
  int p = 0, tot = 0;
  for (int w = 0; w < nw; w++) {
    const int k = s_tmp[w];
    if (w < wave) p += k;
    tot += k;
  }
  __syncthreads();
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int c = i0; c < i1; c += 64) {
    const int i = c + lane;
    const bool b = i < i1 && keep(i);
    const unsigned long long m = __ballot(b);
    if (b) put(i, p + __popcll(m & below));
    p += __popcll(m);
  }
  return tot;
}

// The same over the pairs (k, c) of rows k < nk, columns c in [lo(k),
// hi(k)), in row-major order: each wave a contiguous range of rows, its 64
// lanes over a row's columns.  keep(k, c) tests a pair, put(k, c, p) stores
// it at output position p (from 0).
template <class Lo, class Hi, class Keep, class Put>
__device__ inline int wave_compact_pairs(int nk, Lo lo, Hi hi, Keep keep, Put put, int *s_tmp) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  const int k0 = (int)((long long)nk * wave / nw), k1 = (int)((long long)nk * (wave + 1) / nw);
  int cnt = 0;
  for (int k = k0; k < k1; k++) {
    const int c1 = hi(k);
    for (int c = lo(k) + lane; c - lane < c1; c += 64) cnt += __popcll(__ballot(c < c1 && keep(k, c)));
  }
  if (lane == 0) s_tmp[wave] = cnt;
  __syncthreads();
  int p = 0, tot = 0;
  for (int w = 0; w < nw; w++) {
    const int k = s_tmp[w];
    if (w < wave) p += k;
    tot += k;
  }
  __syncthreads();
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int k = k0; k < k1; k++) {
    const int c1 = hi(k);
    for (int c = lo(k) + lane; c - lane < c1; c += 64) {
      const bool b = c < c1 && keep(k, c);
      const unsigned long long m = __ballot(b);
      if (b) put(k, c, p + __popcll(m & below));
      p += __popcll(m);
    }
  }
  return tot;
}

}  // namespace mdn
